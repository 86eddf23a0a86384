"""CPU-side checks of libnumc_amd.so: exported symbols, the array API
(numc's src/array_core.c behaviour), the error contract of numc_matmul
(src/matmul/dispatch.h:20-59), and the seeded generator against the oracle.
No GEMM is computed here."""
import ctypes as C
import glob
import json
import os
import re

import numpy as np
import pytest

import numc_amd as nc
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "numc_known_answers.json")))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        src = open(h).read()
        for m in re.finditer(r"NUMC_API\s+[\w\s\*]+?\b(numc_\w+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert len(names) > 40
    L = C.CDLL(nc.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # and the Python view binds exactly what the headers declare
    assert set(nc.exported_symbols()) <= names


def test_library_has_gfx950_code():
    blob = open(nc.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_array_basics(ctx):
    a = ctx.array((3, 4), nc.FLOAT32)
    assert a.shape == (3, 4) and a.strides == (16, 4) and a.dtype == nc.FLOAT32
    x = np.arange(12, dtype=np.float32).reshape(3, 4)
    a.write(x)
    assert np.array_equal(a.numpy(), x)
    a.transpose()
    assert a.shape == (4, 3) and a.strides == (4, 16)
    assert np.array_equal(a.numpy(), x.T)
    assert not nc.lib().numc_array_is_contiguous(a.h)
    s = ctx.from_numpy(x).slice(1, 1, 4, 2)
    assert s.shape == (3, 2) and np.array_equal(s.numpy(), x[:, 1:4:2])
    c = nc.Array(nc.lib().numc_array_copy(s.h), ctx)
    assert nc.lib().numc_array_is_contiguous(c.h) and np.array_equal(c.numpy(), x[:, 1:4:2])
    z = ctx.array((2, 5), nc.INT16, zeros=True)
    assert np.array_equal(z.numpy(), np.zeros((2, 5), np.int16))
    v = C.c_int64(-7)
    f = nc.Array(nc.lib().numc_array_fill(ctx.h, (C.c_size_t * 2)(2, 3), 2, nc.INT64, C.byref(v)), ctx)
    assert np.array_equal(f.numpy(), np.full((2, 3), -7, np.int64))


def test_contiguous_and_reshape(ctx):
    x = np.arange(24, dtype=np.int32).reshape(4, 6)
    a = ctx.from_numpy(x).transpose()
    assert nc.lib().numc_array_contiguous(a.h) == 0
    assert a.strides == (16, 4) and np.array_equal(a.numpy(), x.T)
    assert nc.lib().numc_array_reshape(a.h, (C.c_size_t * 3)(2, 3, 4), 3) == 0
    assert np.array_equal(a.numpy(), x.T.reshape(2, 3, 4))
    assert nc.lib().numc_array_reshape(a.h, (C.c_size_t * 1)(5), 1) == -1
    assert nc.last_error()[0] == nc.ERR_SHAPE


def test_checkpoint_restore_keeps_earlier_arrays(ctx):
    x = np.arange(6, dtype=np.float64)
    a = ctx.from_numpy(x)
    cp = nc.lib().numc_ctx_checkpoint(ctx.h)
    for _ in range(5):
        big = ctx.array((1 << 20,), nc.FLOAT64)  # 8 MiB each: forces new blocks
        assert big.shape == (1 << 20,)
    nc.lib().numc_ctx_restore(ctx.h, cp)
    assert np.array_equal(a.numpy(), x)
    b = ctx.from_numpy(x * 2)
    assert np.array_equal(b.numpy(), x * 2) and np.array_equal(a.numpy(), x)


def test_empty_checkpoint_restore():
    c = nc.Ctx()
    cp = nc.lib().numc_ctx_checkpoint(c.h)
    c.array((10,), nc.INT8)
    nc.lib().numc_ctx_restore(c.h, cp)
    a = c.from_numpy(np.ones(3, np.int8))
    assert np.array_equal(a.numpy(), np.ones(3, np.int8))
    c.close()


def test_slice_errors(ctx):
    a = ctx.array((4, 4), nc.FLOAT32)
    s = nc.Slice(2, 0, 1, 1)
    assert not nc.lib().numc_array_slice(a.h, C.byref(s))
    assert nc.last_error()[0] == nc.ERR_BOUNDS
    s = nc.Slice(0, 3, 2, 1)
    assert not nc.lib().numc_array_slice(a.h, C.byref(s))
    assert nc.last_error()[0] == nc.ERR_VALUE


@pytest.mark.parametrize("fn", ["numc_matmul", "numc_matmul_naive"])
def test_error_contract(ctx, fn):
    codes = {"float32": nc.FLOAT32, "float64": nc.FLOAT64}
    for e in GOLD["errors"]:
        arrs = []
        for key, dt in zip(("a", "b", "out"), e["dtypes"]):
            arrs.append(None if e[key] is None else ctx.array(e[key], codes[dt], zeros=True).h)
        got = getattr(nc.lib(), fn)(*arrs)
        assert got == e["code"], e["case"]
        assert nc.last_error()[0] == e["code"]
        assert "matmul" in nc.last_error()[1]


def test_widen_only_with_flag(ctx):
    a = ctx.array((2, 3), nc.INT8, zeros=True)
    b = ctx.array((3, 2), nc.INT8, zeros=True)
    o = ctx.array((2, 2), nc.INT32, zeros=True)
    assert nc.lib().numc_matmul(a.h, b.h, o.h) == nc.ERR_TYPE
    opts = nc.MatmulOpts(0, -1, nc.OPT_WIDEN)
    assert nc.lib().numc_matmul_family(a.h, b.h, o.h, C.byref(opts)) == nc.FAMILY_I8


def test_method_needs_float32(ctx):
    a = ctx.array((2, 2), nc.FLOAT64, zeros=True)
    opts = nc.MatmulOpts(nc.METHOD_SPLIT_BF16, -1, 0)
    assert nc.lib().numc_matmul_family(a.h, a.h, a.h, C.byref(opts)) == nc.ERR_TYPE


def test_no_cpu_fallback_without_gpu(ctx):
    """Without a GPU numc_matmul must fail loudly, never compute on the CPU."""
    import subprocess
    probe = subprocess.run(["python", "-c", "import ctypes;L=ctypes.CDLL('%s');"
                            "import sys;n=ctypes.c_int();sys.exit(0)" % nc.LIB_PATH])
    assert probe.returncode == 0
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    a = ctx.from_numpy(np.ones((4, 4), np.float32))
    o = ctx.array((4, 4), nc.FLOAT32, zeros=True)
    assert nc.lib().numc_matmul(a.h, a.h, o.h) == nc.ERR_INTERNAL
    assert "no HIP device" in nc.last_error()[1]
    assert np.array_equal(o.numpy(), np.zeros((4, 4), np.float32))


@pytest.mark.parametrize("dt,dist,lo,hi", [
    (nc.FLOAT32, nc.DIST_UNIFORM, -1.0, 1.0), (nc.FLOAT64, nc.DIST_UNIFORM, -1.0, 1.0),
    (nc.FLOAT32, nc.DIST_NORMAL, 0, 0), (nc.FLOAT64, nc.DIST_NORMAL, 0, 0),
    (nc.INT32, nc.DIST_INT, -100, 100), (nc.INT8, nc.DIST_INT, -128, 127),
])
def test_generator_matches_oracle(ctx, dt, dist, lo, hi):
    n = 200003
    a = ctx.array((n,), dt)
    a.fill_random(dist, 1234, 5, lo, hi)
    npdt = nc.NP_DTYPES[dt]
    if dist == nc.DIST_UNIFORM:
        want = oracle.fill_uniform(n, npdt, 1234, 5, lo, hi)
    elif dist == nc.DIST_NORMAL:
        want = oracle.fill_normal(n, npdt, 1234, 5)
    else:
        want = oracle.fill_int(n, npdt, 1234, 5, lo, hi)
    assert np.array_equal(a.numpy(), want)


def test_config_tables():
    for fam in range(nc.FAMILY_COUNT):
        n = nc.config_count(fam)
        assert n >= 3
        for i in range(n):
            c = nc.config_get(fam, i)
            assert c["bm"] % 16 == 0 and c["bn"] % 16 == 0 and c["split_k"] >= 1


def test_pointer_entry_error_contract(ctx):
    """numc_amd_gemm validates before touching the GPU: NULL -> dtype -> shape -> strides."""
    L = nc.lib()
    a = np.ones((4, 3), np.float32)
    b = np.ones((3, 5), np.float32)
    c = np.zeros((4, 5), np.float32)
    opts = nc.MatmulOpts(0, -1, 0)

    def call(h, dt, odt, M, N, K, pa, pb, pc, rsa=12, csa=4, rsb=20, csb=4, rso=20, cso=4, o=opts):
        return L.numc_amd_gemm(h, dt, odt, M, N, K, pa, rsa, csa, pb, rsb, csb, pc, rso, cso, C.byref(o))

    pa, pb, pc = a.ctypes.data, b.ctypes.data, c.ctypes.data
    assert call(None, nc.FLOAT32, nc.FLOAT32, 4, 5, 3, pa, pb, pc) == nc.ERR_NULL
    assert call(ctx.h, nc.FLOAT32, nc.FLOAT32, 4, 5, 3, None, pb, pc) == nc.ERR_NULL
    assert call(ctx.h, nc.FLOAT32, nc.FLOAT64, 4, 5, 3, pa, pb, pc) == nc.ERR_TYPE
    assert call(ctx.h, 12, 12, 4, 5, 3, pa, pb, pc) == nc.ERR_TYPE     # past the last dtype
    assert call(ctx.h, nc.INT8, nc.INT32, 4, 5, 3, pa, pb, pc) == nc.ERR_TYPE   # widening needs the flag
    assert call(ctx.h, nc.FLOAT32, nc.FLOAT32, 1 << 30, 5, 3, pa, pb, pc) == nc.ERR_SHAPE
    assert call(ctx.h, nc.FLOAT32, nc.FLOAT32, 4, 5, 3, pa, pb, pc, csa=2) == nc.ERR_VALUE
    bad = nc.MatmulOpts(nc.METHOD_SPLIT_BF16, -1, 0)
    assert call(ctx.h, nc.FLOAT64, nc.FLOAT64, 4, 5, 3, pa, pb, pc, rsa=24, csa=8, rsb=40, csb=8, rso=40, cso=8,
                o=bad) == nc.ERR_TYPE
    assert call(ctx.h, nc.FLOAT32, nc.FLOAT32, 0, 5, 3, pa, pb, pc) == 0      # empty: nothing to do
    if not os.path.exists("/dev/kfd"):
        assert call(ctx.h, nc.FLOAT32, nc.FLOAT32, 4, 5, 3, pa, pb, pc) == nc.ERR_INTERNAL
        assert np.array_equal(c, np.zeros((4, 5), np.float32)), "no CPU fallback"


def test_dtype_header_keeps_numc_generators(tmp_path):
    """include/numc/dtype.h keeps numc's element-type macros, X-macro
    generators and numc_dtype_align: a C file instantiating a table through
    GENERATE_NUMC_TYPES compiles and agrees with numc_dtype_size; where numc's
    sources are present, numc's own src/matmul/kernel.h (which expands
    GENERATE_NUMC_TYPES) compiles against this header too."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    src = tmp_path / "gen.c"
    src.write_text(
        "#include <numc/dtype.h>\n#include <stdio.h>\n"
        "#define ROW(E, T) {E, sizeof(T), NUMC_ALIGNOF(T)},\n"
        "static const struct { int e; size_t s, a; } t[] = {GENERATE_NUMC_TYPES(ROW)};\n"
        "#define ONE(E, T) + 1\n"
        "int main(void) {\n"
        "  int bad = 0, n = sizeof t / sizeof t[0];\n"
        "  for (int i = 0; i < n; i++)\n"
        "    bad += t[i].e != i || t[i].s != numc_dtype_size((NumcDType)i) || t[i].a != numc_dtype_align((NumcDType)i);\n"
        "  bad += (0 GENERATE_INT_NUMC_TYPES(ONE)) != 8 || (0 GENERATE_FLOAT_NUMC_TYPES(ONE)) != 2;\n"
        "  bad += (0 GENERATE_32BIT_NUMC_TYPES(ONE)) != 3 || (0 GENERATE_64BIT_NUMC_TYPES(ONE)) != 3;\n"
        "  printf(\"%d %d\\n\", n, bad);\n  return bad;\n}\n")
    exe = tmp_path / "gen"
    subprocess.run([cc, "-std=gnu11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.split() == ["10", "0"], out.stdout
    ref = "/root/reference"
    clang = "/opt/rocm/lib/llvm/bin/clang"
    if os.path.exists(os.path.join(ref, "src", "matmul", "kernel.h")) and os.path.exists(clang):
        k = tmp_path / "kernel_user.c"
        k.write_text('#include "matmul/kernel.h"\nint main(void) { return 0; }\n')
        r = subprocess.run([clang, "-std=gnu23", "-fsyntax-only", "-fopenmp", "-I", os.path.join(ROOT, "include"),
                            "-I", os.path.join(ref, "src"), str(k)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
