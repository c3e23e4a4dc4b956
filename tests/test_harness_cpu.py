"""The C consumer of include/vxl_sph.h (tests/harness/abi_harness.c) and the
host-sanitizer builds, on the CPU.

* The harness, a plain C program with no Python in the process, gets the same
  lattice as the oracle's restatement of boundary::lattice
  (crates/vxl-phys-fluid/src/boundary.rs:67-88) and the same dam-break scene,
  request footprints and pool limits as the Python binding: the structures
  of the header mean the same from C as from the binding.
* AddressSanitizer + UndefinedBehaviorSanitizer (leak checking on) and
  ThreadSanitizer builds of the library and harness write byte-identical
  output and report nothing; the plate lattice is large enough (64Ki points
  and up) that its volume self-sum runs on several threads.
* The oracle's pin suite and the library's CPU-side suites pass against
  AddressSanitizer + UndefinedBehaviorSanitizer builds loaded into Python
  (the reference runs its own suites under sanitizers in CI,
  .github/workflows/ci.yml:284-328).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import vxl_sph as V
from scenes import bits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "harness")
F = np.float32


def harness(variant, mode, out, env=None, timeout=120):
    exe = os.path.join(HARNESS, "abi_harness" + variant)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C tests/harness all sanitize)")
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe, mode, str(out)], env=e, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"{variant or 'plain'} {mode}: rc {r.returncode}\n{r.stdout}\n{r.stderr[-4000:]}"
    return open(out, "rb").read()


class Reader:
    def __init__(self, raw):
        self.b = raw
        self.o = 0

    def take(self, dtype, count):
        a = np.frombuffer(self.b, dtype, count, self.o)
        self.o += a.nbytes
        return a

    def done(self):
        return self.o == len(self.b)


SPAN = np.dtype([("body", "<u4"), ("origin", "<f4", 3), ("start", "<u4"), ("end", "<u4")])


def parse_cpu(raw):
    r = Reader(raw)
    magic, n, n0, _ = r.take("<u4", 4)
    assert magic == 0x43505531
    vol = r.take("<f4", 1)[0]
    pts = r.take("<f4", 3 * n).reshape(n, 3)
    pn, pn0 = (int(x) for x in r.take("<u4", 2))
    pvol = r.take("<f4", 1)[0]
    sn, snf, sns = r.take("<u4", 3)
    scene = dict(n=sn, nf=snf, pos=r.take("<f4", 3 * sn).reshape(sn, 3), vel=r.take("<f4", 3 * sn).reshape(sn, 3),
                 pmass=r.take("<f4", sn), spans=r.take(SPAN, sns))
    reqs = r.take("<u8", 18).reshape(6, 3)
    slots, cells = r.take("<u8", 2)
    rb = r.take("<u4", 1)[0]
    assert r.done()
    return dict(lattice=(pts, int(n0), vol), plate=(pn, pn0, pvol), scene=scene, reqs=reqs, limits=(int(slots), int(cells), int(rb)))


def test_harness_cpu_agrees_with_oracle_and_binding(tmp_path):
    out = parse_cpu(harness("", "cpu", tmp_path / "cpu.bin"))
    pts, n0, vol = out["lattice"]
    opts, on0, ovol = O.lattice_box((0.2, 0.05, 0.15), F(0.02), F(0.04), 2)
    np.testing.assert_array_equal(bits(pts), bits(opts))
    assert n0 == on0 and bits(np.array([vol])) == bits(np.array([ovol]))

    pn, pn0, pvol = out["plate"]
    opts, on0, ovol = O.lattice_box((1.0, 0.02, 1.0), F(0.01), F(0.02), 1)
    assert (pn, pn0) == (len(opts), on0) and bits(np.array([pvol])) == bits(np.array([ovol]))

    cfg = V.dam_break_config(0.02)
    data = V.dam_break(cfg, (0.4, 0.3, 0.3), (9, 8, 7), 0.02, seed=11)
    sc = out["scene"]
    assert (sc["n"], sc["nf"]) == (data.n, data.n_fluid)
    for k in ("pos", "vel", "pmass"):
        np.testing.assert_array_equal(bits(sc[k]), bits(getattr(data, k)), err_msg=k)
    assert [(int(s["body"]), tuple(F(s["origin"])), int(s["start"]), int(s["end"])) for s in sc["spans"]] == \
        [(b, tuple(F(o)), s0, s1) for b, o, s0, s1 in data.spans]

    reqs = V.sample_requests(3, 6, 20.0, 4096, 65536, 10, 40, 4)
    want = [(r.steps, V.request_bytes(r), r.nboxes and 5 + r.nboxes) for r in reqs]
    assert [tuple(int(x) for x in row) for row in out["reqs"]] == [tuple(int(x) for x in w) for w in want]
    assert out["limits"] == V.pool_limits()


def _sanitizer_runtime(name):
    clang = "/opt/rocm/llvm/bin/clang"
    if not os.path.exists(clang):
        pytest.skip("clang of the ROCm toolchain is absent")
    rd = subprocess.run([clang, "-print-resource-dir"], capture_output=True, text=True).stdout.strip()
    rt = os.path.join(rd, "lib", "linux", f"libclang_rt.{name}-x86_64.so")
    if not os.path.exists(rt):
        pytest.skip(f"{rt} is absent")
    return rt


def test_harness_cpu_under_asan_ubsan_and_tsan(tmp_path):
    """Leaks are checked here (the CPU path creates and frees everything it uses)."""
    plain = harness("", "cpu", tmp_path / "plain.bin")
    asan = harness("_asan", "cpu", tmp_path / "asan.bin",
                   env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:halt_on_error=1",
                        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})
    tsan = harness("_tsan", "cpu", tmp_path / "tsan.bin", env={"TSAN_OPTIONS": "halt_on_error=1"})
    assert asan == plain and tsan == plain


def instrumented(lib, prefix):
    """The library calls into the sanitizer runtime (it was built instrumented)."""
    syms = subprocess.run(["nm", "-D", "--undefined-only", lib], capture_output=True, text=True).stdout
    return any(line.split()[-1].startswith(prefix) for line in syms.splitlines() if line.strip())


def _pytest_under(preload, env, files, timeout):
    e = dict(os.environ)
    e.update(env)
    e["LD_PRELOAD"] = preload
    e["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1"
    e["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"]
                       + files, cwd=ROOT, env=e, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    return r.stdout


def test_oracle_pins_under_asan_ubsan():
    lib = os.path.join(ROOT, "oracle", "liboracle_asan.so")
    if not os.path.exists(lib):
        pytest.fail(f"{lib} is not built (make -C oracle sanitize)")
    rt = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(rt):
        pytest.skip("the compiler's libasan is absent")
    assert instrumented(lib, "__asan_report_") and instrumented(lib, "__ubsan_handle_")
    out = _pytest_under(rt, {"ORACLE_LIB": lib}, ["tests/test_oracle.py"], 600)
    assert " passed" in out


def test_library_cpu_suites_under_asan_ubsan():
    lib = os.path.join(ROOT, "rust-wl_amd", "libvxl_sph_asan.so")
    if not os.path.exists(lib):
        pytest.fail(f"{lib} is not built (make -C rust-wl_amd sanitize)")
    assert instrumented(lib, "__asan_report_") and instrumented(lib, "__ubsan_handle_")
    assert instrumented(os.path.join(ROOT, "rust-wl_amd", "libvxl_sph_tsan.so"), "__tsan_")
    out = _pytest_under(_sanitizer_runtime("asan"), {"VXL_SPH_LIB": lib},
                        ["tests/test_abi_cpu.py", "tests/test_vxld.py"], 300)
    assert " passed" in out
