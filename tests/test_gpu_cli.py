"""The drop-in CLIs run end to end on the GPU.

Spinless: the README runs `./blqh -n 8 -e 8 -u 4 -j 4 -k 4 -t 4 -d <d>`
(README.md:58-75 for d = 0.0, README.md:85-102 for d = 3.0; report format of
BLQH_without_spin/main.cpp:25-71).  Compared line by line: the two headers,
`nHilbert: =79`, the ED energy at its printed 6 digits, the Lanczos energy at
its printed 10 digits, the wave-function listing's amplitude column (sorted
magnitudes are gauge invariant) and the shape of every listing line.  The
basis indices of the listing are not compared: the reference's representative
of each orbit depends on its libstdc++'s hash order, and this libstdc++ does
not reproduce the README's indices.  Each listed index must instead name the
basis word its bit columns show.

Spinful: `blqh_spin` (BLQH_with_spin/main.cpp:25-99) on a small sector: its
three-line listing blocks, the energies against the library's own dense and
Lanczos results at the printed 10 digits."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import blqh_amd

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "readme_runs.json")
LINE = re.compile(r"^ *(\d+) :   \|((?: +(?:\d+|_)){8})\)\|((?: +(?:\d+|_)){8})\)   ([01]{24})  (\S+)$")


def run(exe, args, env=None):
    print("run", os.path.basename(exe), " ".join(args), flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120,
                       env=None if env is None else dict(os.environ, **env))
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.splitlines()


def occ_word(cols, shift):
    w = 0
    for k, tok in enumerate(cols.split()):
        if tok != "_":
            assert int(tok) == k, cols
            w |= 1 << (k + shift)
    return w


@pytest.mark.parametrize("key", ["d0", "d3"])
def test_readme_run(key):
    g = json.load(open(GOLD))[key]
    out = run(blqh_amd.CLI_PATH, ["-n", "8", "-e", "8", "-u", "4", "-j", "4", "-k", "4", "-t", "4", "-d",
                                  "%.1f" % g["d"], "-s", "1"])
    assert out[0] == "----------- ED results --------------"
    assert out[1] == "nHilbert: =%d" % g["nHilbert"]
    assert out[2] == "E_gs:= %.6g" % g["ed_e_per_el"]
    assert out[3] == "----------- Lanczos results ---------"
    assert out[4] == "E_gs:= %.10g" % g["lanczos_e_per_el"]
    assert out[5] == "# ground state wave function"
    B = blqh_amd.Basis(8, 8, 4, 4, 4)
    rows = [LINE.match(s) for s in out[6:]]
    assert all(rows), out[6:]
    amps = [float(m.group(5)) for m in rows]
    want = g["amplitudes"]
    assert ["%.10g" % a for a in amps[:len(want)]] == ["%.10g" % a for a in want]
    assert len(rows) == min(11, len(want) + 1)
    for m in rows:
        idx = int(m.group(1))
        word = occ_word(m.group(2), 0) | occ_word(m.group(3), 8)
        assert int(m.group(4), 2) == word
        assert int(B.id[idx]) == word
    assert amps == sorted(amps, reverse=True)


def test_spinful_run(monkeypatch):
    nphi, nel, nu, nsu, J, K = 4, 4, 2, 2, 0, 1
    d, dsas = 1.0, 0.05
    # The reference's fixed 200 steps run far past this small sector's
    # dimension: without reorthogonalisation the printed Rayleigh quotient
    # then depends on the last bits of every step, i.e. on the SpMV's
    # summation order.  Both runs use the same SpMV variant.
    plan = {"BLQH_SPMV_PLAN": "0,16,256,0,0"}
    out = run(blqh_amd.CLI_PATH + "_spin", ["-n", str(nphi), "-e", str(nel), "-U", str(nu), "-u", str(nsu), "-J",
                                            str(J), "-K", str(K), "-d", str(d), "-S", str(dsas), "-s", "3"], env=plan)
    monkeypatch.setenv("BLQH_SPMV_PLAN", plan["BLQH_SPMV_PLAN"])
    B = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=4, nel_su=nsu)
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d, delta_sas=dsas))
    e0, ep, _ = H.dense_ground_state()
    r = H.lanczos(lam=200, seed=3, tol=0)  # the CLI runs the reference's fixed 200 steps
    assert out[0] == "----------- ED results --------------"
    assert out[1] == "nHilbert: =%d" % B.n
    assert out[2] == "E_gs:= %.10g" % (ep / nel)
    assert out[3] == "----------- Lanczos results ---------"
    assert out[4] == "E_gs:= %.10g" % (r["E_rq"] / nel)
    body = out[5:]
    assert len(body) % 3 == 0 and body
    head = re.compile(r"^ {9}\((\d+)\): ([01]{32})  (\S+)$")
    up = re.compile(r"^upper layer:   \|((?: +(?:\d+|_)){4})\)\|((?: +(?:\d+|_)){4})\)\|$")
    down = re.compile(r"^ down layer:   \|((?: +(?:\d+|_)){4})\)\|((?: +(?:\d+|_)){4})\)   $")
    amps = []
    for i in range(0, len(body), 3):
        h, u, w = head.match(body[i]), up.match(body[i + 1]), down.match(body[i + 2])
        assert h and u and w, body[i:i + 3]
        idx = int(h.group(1))
        word = int(B.id[idx])
        assert int(h.group(2), 2) == word
        # upper-layer species print their orbital index, lower-layer species
        # their bit position (BLQH_with_spin/main.cpp:66-92)
        assert occ_word(u.group(1), 0) | occ_word(u.group(2), nphi) == word & ((1 << 2 * nphi) - 1)
        low = [tok for tok in (w.group(1) + w.group(2)).split()]
        bits = [k + 2 * nphi for k, tok in enumerate(low) if tok != "_"]
        assert all(int(tok) == k + 2 * nphi for k, tok in enumerate(low) if tok != "_")
        assert sum(1 << b for b in bits) == word & ~((1 << 2 * nphi) - 1)
        amps.append(float(h.group(3)))
        assert float(h.group(3)) == pytest.approx(abs(r["psi"][idx]), rel=1e-6, abs=1e-9)
    assert amps == sorted(amps, reverse=True)
    assert len(amps) <= 11
