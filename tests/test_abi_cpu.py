"""CPU-only checks of the drop-in boundary: the C-ABI library loads, exports
every entry point include/splat_gs.h declares, and its structs match the
ctypes mirrors byte for byte. No GPU calls are made."""
import ctypes as C
import os
import re
import subprocess

import pytest

import gsabi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "splat_gs.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = gsabi.lib()
    declared = _declared()
    assert "gs_setup" in declared and "gs_step" in declared
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) == set(gsabi.SIGNATURES), set(declared) ^ set(gsabi.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", gsabi.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    for s in declared:
        assert re.search(rf"\bT {s}\b", out), s


def test_default_opts_are_the_references():
    o = gsabi.default_opts()
    assert o.size == C.sizeof(gsabi.GsOpts)
    assert o.eCut == pytest.approx(4.5) and o.aMin == pytest.approx(0.0039)
    assert o.radClamp == 1.0 and o.shDeg == 2
    assert o.anisoReg == pytest.approx(0.005) and o.opacityReg == pytest.approx(0.01)
    assert o.shLr == pytest.approx(1.25e-4) and o.minScale == pytest.approx(1e-4)
    # maxSplats unset: capacity ceiling 600000 and the 12M key budget (trainer.js:48-49, :105-107)
    assert o.maxIters == 60000 and o.maxSplats == 0 and o.entriesCap == 0 and o.capMult == 4.0
    assert o.loss == gsabi.GS_LOSS_CHARBONNIER and o.gradAccum == gsabi.GS_ACCUM_FIXED
    assert o.camOpt == 0 and o.expComp == 0          # trainer.js:468, :511: both off by default
    assert o.tile == 16 and o.batch == 256             # shaders.js:32 TILE


def test_struct_layouts_match_the_header(tmp_path):
    """Compile a probe against the header and compare sizes/offsets."""
    probe = tmp_path / "probe.c"
    fields = {
        "gs_opts": [f[0] for f in gsabi.GsOpts._fields_],
        "gs_camera": [f[0] for f in gsabi.GsCamera._fields_],
        "gs_image": [f[0] for f in gsabi.GsImage._fields_],
        "gs_info": [f[0] for f in gsabi.GsInfo._fields_],
        "gs_debug_out": [f[0] for f in gsabi.GsDebugOut._fields_],
        "gs_kernel_time": [f[0] for f in gsabi.GsKernelTime._fields_],
        "gs_refine_opts": [f[0] for f in gsabi.GsRefineOpts._fields_],
        "gs_refine_result": [f[0] for f in gsabi.GsRefineResult._fields_],
        "gs_cam_state": [f[0] for f in gsabi.GsCamState._fields_],
        "gs_step_stats": [f[0] for f in gsabi.GsStepStats._fields_],
        "gs_readback_result": [f[0] for f in gsabi.GsReadbackResult._fields_],
        "gs_tune_entry": [f[0] for f in gsabi.GsTuneEntry._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for s, fs in fields.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-o", str(exe), str(probe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    py = {"gs_opts": gsabi.GsOpts, "gs_camera": gsabi.GsCamera, "gs_image": gsabi.GsImage,
          "gs_info": gsabi.GsInfo, "gs_debug_out": gsabi.GsDebugOut, "gs_kernel_time": gsabi.GsKernelTime,
          "gs_refine_opts": gsabi.GsRefineOpts, "gs_refine_result": gsabi.GsRefineResult,
          "gs_cam_state": gsabi.GsCamState, "gs_step_stats": gsabi.GsStepStats,
          "gs_readback_result": gsabi.GsReadbackResult, "gs_tune_entry": gsabi.GsTuneEntry}
    for s, cls in py.items():
        assert int(got[s]) == C.sizeof(cls), s
        for f, _ in cls._fields_:
            assert int(got[f"{s}.{f}"]) == getattr(cls, f).offset, (s, f)


def test_library_is_built_for_gfx950():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-n", gsabi.LIB_PATH], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True).stdout
    blob = open(gsabi.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_create_rejects_ecut_outside_supported_range():
    """The rasterizer's branch-free exp is exact only for exponents in
    [-80, 0]; options outside that range are refused before any HIP call."""
    L = gsabi.lib()
    for bad in (80.5, -1.0, float("nan"), float("inf")):
        o = gsabi.default_opts()
        o.eCut = bad
        h = C.c_void_p()
        assert L.gs_create(C.byref(o), C.byref(h)) == -1, bad  # GS_E_ARG
        assert not h.value


MATH_PROBE = r"""
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include "gs_math.h"
int main(void) {
  /* every float in [-87, 0] (sign bit set, magnitude 0 .. 87): ulp distance
     of gs_expf and gs_expf_raster from the correctly rounded e^x (libm's
     double exp rounded once to float; e^-87 is still a normal float) */
  const uint32_t top = gs_f2u(87.0f);
  unsigned long long n = 0, hp[4] = {0, 0, 0, 0}, hr[4] = {0, 0, 0, 0};
#pragma omp parallel
  {
    unsigned long long cn = 0, cp[4] = {0, 0, 0, 0}, cr[4] = {0, 0, 0, 0};
#pragma omp for schedule(static)
    for (long long mm = 0; mm <= (long long)top; mm++) {
      const float x = gs_u2f(0x80000000u | (uint32_t)mm);
      const long long t = (long long)gs_f2u((float)exp((double)x));
      const long long dp = llabs((long long)gs_f2u(gs_expf(x)) - t);
      const long long dr = llabs((long long)gs_f2u(gs_expf_raster(x)) - t);
      cn++;
      cp[dp > 3 ? 3 : dp]++;
      cr[dr > 3 ? 3 : dr]++;
    }
#pragma omp critical
    {
      n += cn;
      for (int i = 0; i < 4; i++) { hp[i] += cp[i]; hr[i] += cr[i]; }
    }
  }
  /* gs_expf_mid == gs_expf bit for bit on every float of [-87, 87];
     rsub: gs_expf_raster results below FLT_MIN (its final p 2^k is exact,
     and equal to the device's ldexp, only while the result is normal) */
  unsigned long long midbad = 0, rsub = 0;
#pragma omp parallel for reduction(+ : midbad, rsub) schedule(static)
  for (long long mm = 0; mm <= (long long)top; mm++) {
    const float xn = gs_u2f(0x80000000u | (uint32_t)mm), xp = gs_u2f((uint32_t)mm);
    midbad += gs_f2u(gs_expf_mid(xn)) != gs_f2u(gs_expf(xn));
    midbad += gs_f2u(gs_expf_mid(xp)) != gs_f2u(gs_expf(xp));
    rsub += gs_f2u(gs_expf_raster(xn)) < 0x00800000u;
  }
  printf("%llu %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu\n", n, hp[0], hp[1], hp[2], hp[3], hr[0], hr[1],
         hr[2], hr[3], midbad, rsub);
  return 0;
}
"""


def test_exps_against_correctly_rounded_exp_exhaustively(tmp_path):
    """On every float of [-87, 0]: gs_expf (projection, sigmoid) is within 1
    ulp of the correctly rounded e^x, and gs_expf_raster (the rasterizers'
    alpha, argument clamped to [-87, 0] by the caller) within 2 ulp, with the
    2-ulp cases rare; WGSL allows exp 3 + 2|x| ulp. gs_expf_mid (scales and
    sigmoids, branch-free) equals gs_expf bit for bit on every float of
    [-87, 87]. Every gs_expf_raster result on [-87, 0] is a normal float, so
    its final scaling by 2^k is exact (the device does it with ldexp)."""
    src = tmp_path / "m.c"
    src.write_text(MATH_PROBE)
    exe = tmp_path / "m"
    flags = ["-O2", "-ffp-contract=off", "-fopenmp"]
    cpu = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    if re.search(r"\bfma\b", cpu) and re.search(r"\bsse4_1\b", cpu):
        flags += ["-mfma", "-msse4.1"]   # hardware fmaf/rintf: same IEEE results, 20x faster than libm
    subprocess.check_call(["gcc", *flags, "-I", os.path.join(ROOT, "splat.js_amd", "csrc"),
                           "-o", str(exe), str(src), "-lm"])
    n, p0, p1, p2, p3, r0, r1, r2, r3, midbad, rsub = map(int, subprocess.check_output([str(exe)], text=True).split())
    assert n == 0x42ae0000 + 1 and midbad == 0
    assert rsub == 0   # every gs_expf_raster result on [-87, 0] is a normal float
    assert p2 == 0 and p3 == 0 and p1 < n // 100
    assert r3 == 0 and r2 < n // 50 and r0 > n * 9 // 10
