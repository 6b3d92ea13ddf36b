"""The gfx950 code of libdclust.so evaluates the FP64 neighbour predicate
without fused multiply-adds.

The oracle (cpu-dbscan/dbscan_rtree_exp.cpp:351-358) rounds every product and
every sum: fl(fl(fl(dx0^2) + fl(dx1^2)) [+ fl(dx2^2)]).  A contracted
fma(dx1, dx1, fl(dx0^2)) rounds once less and can flip a pair at d^2 = eps^2
(tests/test_gpu_boundary.py feeds such pairs to the GPU).  The HIP sources
are built with -ffp-contract=off; this reads the AMD assembly of the library
actually built here and checks that every FP64 FMA belongs to an IEEE
division expansion (v_div_scale .. v_div_fixup, five FMAs each), which the
grid set-up's 1/w and extent/r divisions use and which rounds correctly.
"""
import os
import re
import subprocess

import pytest

import dclust

LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
FMA64 = re.compile(r"\bv_(fma|fmac|mad|mac)_f64|\bv_pk_fma_f64")


def _code_objects(lib, tmp_path):
    """gfx950 code objects of every translation unit linked into `lib`."""
    fat = tmp_path / "fatbin"
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", lib, str(fat)],
                   check=True)
    blob = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), blob)]
    assert starts, "no offload bundle in .hip_fatbin"
    cos = []
    for k, s in enumerate(starts):
        e = starts[k + 1] if k + 1 < len(starts) else len(blob)
        part = tmp_path / f"b{k}.bundle"
        part.write_bytes(blob[s:e])
        listed = subprocess.run([f"{LLVM}/clang-offload-bundler", "--list", "--type=o",
                                 f"--input={part}"], capture_output=True, text=True).stdout
        tgt = [t for t in listed.split() if "gfx950" in t]
        if not tgt:
            continue
        co = tmp_path / f"b{k}.co"
        subprocess.run([f"{LLVM}/clang-offload-bundler", "--unbundle", "--type=o",
                        f"--input={part}", f"--targets={tgt[0]}", f"--output={co}"], check=True)
        cos.append(co)
    return cos


def _functions(asm):
    """{symbol: [instruction lines]} of a disassembly."""
    out, cur = {}, None
    for line in asm.splitlines():
        m = re.match(r"^[0-9a-f]+ <(.+)>:$", line)
        if m:
            cur = m.group(1)
            out[cur] = []
        elif cur and line.startswith("\t"):
            out[cur].append(line.strip())
    return out


@pytest.mark.skipif(not os.path.exists(f"{LLVM}/llvm-objdump"), reason="ROCm LLVM tools absent")
def test_no_fp64_fma_outside_division(tmp_path):
    cos = _code_objects(dclust.LIB_PATH, tmp_path)
    assert len(cos) >= 3  # api, index, cluster, one-step translation units
    names = []
    stray = []
    for co in cos:
        asm = subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--mcpu=gfx950", str(co)],
                             capture_output=True, text=True, check=True).stdout
        for name, ins in _functions(asm).items():
            names.append(name)
            ops = [i.split()[0] for i in ins if i.split()]
            # fp64 -> 64-bit integer conversion: lo = x - floor(x * 2^-32) * 2^32,
            # an exact fma with the literal -2^32 (0xc1f00000)
            conv = sum(1 for i in ins if FMA64.match(i) and "0xc1f00000" in i)
            fma = sum(1 for op in ops if FMA64.match(op)) - conv
            div = sum(1 for op in ops if op.startswith("v_div_fixup_f64"))
            # a correctly rounded FP64 division expands to two Newton steps on
            # the reciprocal (fma + fmac each) and one residual fma, then
            # v_div_fmas / v_div_fixup: 5 FMAs per division and no others
            if fma != 5 * div:
                stray.append((name, fma, div))
    assert len(names) > 30
    assert not stray, stray[:5]
    # the predicate kernels are among the functions checked
    joined = " ".join(names)
    for k in ("k_core", "k_adj", "k_far", "k_border", "k_os_step"):
        assert k in joined, k
