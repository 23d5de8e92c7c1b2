"""CPU sanitizer runs of the host-side code.

1. The oracle built with -fsanitize=address,undefined (oracle/Makefile `sanitize`) runs
   golden and synthetic workloads that reach every oracle function: the explicit and the
   predictor-corrector slice loop (deposition, Poisson, multigrid, push, the PC mixing),
   orders 0-3, the beam push over several steps with external fields, SALAME, tile keys.
2. The C ABI's host code (libhpq.so built with the sanitizers on the host side only,
   hipcc -Xarch_host -fsanitize=...) through the paths that run without a GPU: argument
   validation of hpq_create and of every entry point given a null context, the error of a
   context whose device set-up fails, and hpq_destroy of such a partial context.

Each runs in a subprocess with libasan preloaded (Python itself is not instrumented) and
must end normally without a sanitizer report.
"""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def _clang_asan():
    import glob
    got = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return got[-1] if got else None


def _run(script, preload, timeout=600):
    """Runs `script` with the sanitizer runtimes `preload` (gcc's libasan + libubsan for the
    gcc-built oracle, clang's ASan runtime, which carries the UBSan handlers, for hipcc)."""
    if not preload or not all(preload):
        pytest.skip("sanitizer runtime not found")
    env = dict(os.environ)
    env["LD_PRELOAD"] = ":".join(preload)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["PYTHONPATH"] = ROOT + os.pathsep + os.path.join(ROOT, "tests")
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(script)], env=env, capture_output=True, text=True,
                       timeout=timeout, cwd=ROOT)
    report = r.stdout[-4000:] + r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, report
    assert r.returncode == 0, report
    assert "sanitized run done" in r.stdout, report


def test_oracle_under_address_and_undefined_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8", "sanitize"], check=True,
                   capture_output=True, timeout=600)
    _run("""
        import numpy as np
        from oracle import oracle
        oracle.use_library("oracle/lib/liboracle_san.so")
        from oracle.oracle import OracleSim, tile_keys
        from helpers import deck_run, golden_cases, run_golden, small_config, salame_case, random_plasma
        from hipace2_amd.inputs import BeamSpec, make_beam
        cases = golden_cases()
        for name in ("linear_wake.normalized.1Rank", "blowout_wake_explicit.2Rank", "beam_evolution.1Rank"):
            cfg, beam, max_step = deck_run(cases[name])
            cfg.nz = min(cfg.nz, 24)
            orc = OracleSim(cfg, beam, nthreads=2)
            run_golden(orc, cfg, min(max_step, 2))
            orc.checksums()
        for order in (0, 1, 2, 3):
            for explicit in (True, False):
                cfg = small_config(n=31, nz=8, order=order, dtype=2 if order else 1, explicit=explicit)
                beam = make_beam(cfg, BeamSpec(injection_type="fixed_weight", profile="gaussian", density=3.0,
                                               position_std=(0.3, 0.3, 1.0), num_particles=3000,
                                               u_mean=(0.0, 0.0, 1000.0), seed=order))
                orc = OracleSim(cfg, beam, nthreads=2)
                ez = orc.capture("Ez")
                orc.run_step()
                assert np.isfinite(ez).all()
                d = random_plasma(cfg, 2000, seed=order)
                keys, nb = tile_keys(cfg, d["x"], d["y"], d["valid"], 8)
                assert keys.max() < nb
        cfg, beam = salame_case(nz=24)
        orc = OracleSim(cfg, beam, nthreads=2)
        orc.enable_salame(beam.salame)
        orc.run_step()
        print("sanitized run done")
    """, [_runtime("libasan.so"), _runtime("libubsan.so")])


def test_host_abi_under_address_and_undefined_sanitizers(tmp_path):
    from hipace2_amd import build as hb
    lib = str(tmp_path / "libhpq_san.so")
    hb.build_sanitized(lib, str(tmp_path / "obj"))
    _run(f"""
        import ctypes as C
        from hipace2_amd import _abi
        _abi.LIB_PATH = {lib!r}
        L = _abi.lib()
        p = _abi.Params()
        out = C.c_void_p()
        assert L.hpq_create(None, C.byref(out)) == 1
        # invalid parameters are refused before any device call
        for field, value in (("nx", 2), ("depos_order", 7), ("precision", 9), ("derivative_type", 5),
                             ("deposition", 3), ("tile_size", -1)):
            q = _abi.Params()
            q.nx, q.ny, q.nz = 63, 63, 8
            q.xlo, q.xhi, q.ylo, q.yhi, q.zlo, q.zhi = -1.0, 1.0, -1.0, 1.0, -1.0, 1.0
            q.depos_order, q.derivative_type = 2, 2
            setattr(q, field, value)
            rc = L.hpq_create(C.byref(q), C.byref(out))
            assert rc in (1, 5), (field, rc)
            assert L.hpq_last_error(out)
            L.hpq_destroy(out)
        # a valid set-up with no GPU: the context reports the device error and is destroyed
        q = _abi.Params()
        q.nx, q.ny, q.nz = 63, 63, 8
        q.xlo, q.xhi, q.ylo, q.yhi, q.zlo, q.zhi = -1.0, 1.0, -1.0, 1.0, -1.0, 1.0
        q.depos_order, q.derivative_type, q.explicit_solver = 2, 2, 1
        rc = L.hpq_create(C.byref(q), C.byref(out))
        assert rc != 0 and L.hpq_last_error(out)
        L.hpq_destroy(out)
        # every entry point refuses a null context
        for name in ("hpq_deposit_current", "hpq_explicit_deposition", "hpq_advance_plasma_particles",
                     "hpq_poisson_dirichlet", "hpq_solve_poisson_psi_exmby_eypbx_ez_bz", "hpq_mg_solve1",
                     "hpq_begin_step", "hpq_run_slices", "hpq_bin_tiles", "hpq_bin_tiles_noalloc",
                     "hpq_adjoint_backward", "hpq_set_salame"):
            f = getattr(L, name)
            args = [0.0 if t is C.c_double else 0 if t in (C.c_int, C.c_int64) else None for t in f.argtypes]
            assert f(*args) != 0, name
        print("sanitized run done")
    """, [_clang_asan()])
