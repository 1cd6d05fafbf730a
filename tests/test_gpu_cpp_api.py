"""The header-only C++ API (include/varrcwa.hpp), driven by a compiled C++ program the
way a caller of the reference's classes drives them, against the CPU oracle.

tests/cpp/cpp_api_check is built by varrcwa_amd.build (host-only g++, linked against
libvarrcwa.so). Tolerance: relative 1e-9 on the basis-invariant S transforms.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "cpp_api_check"
TOL = 1e-9


def run_cpp(tmp_path, scen, method, nh, param):
    if not EXE.exists():
        raise FileNotFoundError(f"{EXE} is not built; run `python -m varrcwa_amd.build`")
    tag = "".join(ch if ch.isalnum() else "_" for ch in scen)[-40:]
    out = tmp_path / f"{tag}_{method}.bin"
    print(f"cpp_api_check {scen} {method} nh={nh} param={param}", flush=True)
    r = subprocess.run([str(EXE), scen, method, str(nh), repr(param), str(out)], capture_output=True, text=True,
                       timeout=100)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    head = np.frombuffer(raw[:12], dtype=np.int32)
    n_used, n_eig, N = (int(v) for v in head)
    data = np.frombuffer(raw[12:], dtype=np.complex128)
    mats, off = [], 0
    for k in range(10):
        cnt = N if k in (6, 9) else N * N
        v = data[off:off + cnt]
        off += cnt
        mats.append(v.copy() if k in (6, 9) else v.reshape(N, N, order="F").copy())
    assert off == data.size
    return O.SResult(*mats, N=n_used, nEigVal=n_eig)


def close(got, ref, tol=TOL):
    a, b = O.invariant_transforms(got), O.invariant_transforms(ref)
    for k in a:
        err = np.abs(a[k] - b[k]).max() / np.abs(b[k]).max()
        assert err < tol, (k, err)


@pytest.mark.gpu
def test_cpp_rcwa_metasurface(tmp_path):
    g = G.metasurface(2)
    ref = O.rcwa_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 6)
    close(run_cpp(tmp_path, "metasurface", "rcwa", 2, 6), ref)


@pytest.mark.gpu
def test_cpp_rcwa_var12_pml(tmp_path):
    g = G.var12(2, 2)
    ref = O.rcwa_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 5)
    close(run_cpp(tmp_path, "var12", "rcwa", 2, 5), ref)


@pytest.mark.gpu
def test_cpp_redheffer_metasurface(tmp_path):
    g = G.metasurface(2)
    pe = 5e-2
    ref = O.redheffer_integrate(G.oracle_sampler(g), g["z0"], g["z1"], pe)
    got = run_cpp(tmp_path, "metasurface", "redheffer", 2, pe)
    assert got.N == ref.N
    close(got, ref)


@pytest.mark.gpu
def test_cpp_redheffer_compute_many(tmp_path):
    """RedhefferIntegrator::computeMany over two integrators of one layer: the same N and
    S-matrix as the oracle's single-integrator refinement."""
    g = G.metasurface(2)
    pe = 5e-2
    ref = O.redheffer_integrate(G.oracle_sampler(g), g["z0"], g["z1"], pe)
    got = run_cpp(tmp_path, "metasurface", "redheffermany", 2, pe)
    assert got.N == ref.N
    close(got, ref)


@pytest.mark.gpu
def test_cpp_differential_metasurface(tmp_path):
    g = G.metasurface(1)
    ref = O.differential_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 32)
    close(run_cpp(tmp_path, "metasurface", "differential", 1, 32), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("scen", ["user", "usergpu"])
def test_cpp_user_sampler_subclass(tmp_path, scen):
    """The sampler plugin boundary (src/core/LayerSampler.h:5-77): a LayerSampler subclass
    written by the caller (overriding sample(), or sampleOnGPU()) drives all three
    integrators and matches the oracle on the metasurface scenario."""
    g = G.metasurface(2)
    ref = O.rcwa_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 6)
    close(run_cpp(tmp_path, scen, "rcwa", 2, 6), ref)
    pe = 5e-2
    ref = O.redheffer_integrate(G.oracle_sampler(g), g["z0"], g["z1"], pe)
    got = run_cpp(tmp_path, scen, "redheffer", 2, pe)
    assert got.N == ref.N
    close(got, ref)
    g1 = G.metasurface(1)
    ref = O.differential_integrate(G.oracle_sampler(g1), g1["z0"], g1["z1"], 16)
    close(run_cpp(tmp_path, scen, "differential", 1, 16), ref)


@pytest.mark.gpu
def test_cpp_rcwa_gds_contour(tmp_path):
    """GDSIISampler through the C++ classes on the reference's layout file
    (src/tests/test_gds_cpu.cpp:129-145 at 5 x 3 harmonics, 10 slices)."""
    f = ROOT / "tests" / "golden" / "contour_extract.gds"
    o = O.GDSIISampler(1.55, 5, 3, str(f), shift=(2.3647, 0.0, 0.0), lPML=0.5, rPML=0.5, lspace=1.0, rspace=1.3)
    ref = O.rcwa_integrate(o, -1.63, 1.36, 10)
    close(run_cpp(tmp_path, f"gds:{f}", "rcwa", 2, 10), ref)


def test_cpp_header_compiles():
    """The header compiles as C++17 with the host compiler alone (no HIP headers)."""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-I", str(ROOT / "include"), "-"],
                       input='#include "varrcwa.hpp"\nint main(){return 0;}\n', capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
