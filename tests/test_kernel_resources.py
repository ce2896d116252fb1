"""Register, scratch and occupancy budgets of the stage kernels, read from the built library.

The stage kernels' speed rests on their register budgets (waves per SIMD) and on staying free of
scratch spills, which depend on the compiler's scheduling (the kernels pin parts of it with empty
asm statements and scheduling barriers).  This reads the AMDGPU code-object metadata of this
repository's own build of fve_dg_amd/libfvdg.so (the .hip_fatbin offload bundles, unbundled for
gfx950) and checks each stage kernel of the benchmark configurations against its budget, so that
a compiler or source change that costs occupancy or adds spills fails here, on the CPU.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

LIB = Path(__file__).resolve().parents[1] / "fve_dg_amd" / "libfvdg.so"
LLVM = Path("/opt/rocm/llvm/bin")
TOOLS = [LLVM / "llvm-objcopy", LLVM / "clang-offload-bundler", LLVM / "llvm-readobj"]
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"

# kernel (as c++filt prints it, spaces removed) -> (min resident wavefronts per compute unit, from
# registers and LDS; max scratch bytes per lane)
BUDGET = {
    # P3 (C2, C3): FP64 drop-in, FP32 face tiles, BF16x3 on the face tiles
    "k_face_flux_dpp<3,double,256>": (12, 0),
    "k_elem_mfma<3,double,true,false,false>": (8, 0),
    "k_face_tile<3,float>": (8, 0),
    "k_elem_mfma<3,float,true,true,false>": (12, 0),
    "k_elem_mfma<3,float,true,true,true>": (8, 0),
    "k_face_tile_bf3<3>": (8, 0),
    # P4 (C5): 23 KB element and face-tile tiles hold six wavefronts per compute unit; the FP64
    # element kernel spills at its 256-register cap
    "k_face_flux_dpp<4,double,256>": (8, 0),  # 56 KB of tables per block
    "k_elem_mfma<4,double,true,false,false>": (6, 252),
    "k_face_tile<4,float>": (6, 0),
    "k_elem_mfma<4,float,true,true,false>": (6, 0),
    "k_elem_mfma<4,float,true,true,true>": (6, 0),
    "k_face_tile_bf3<4>": (6, 0),
    # P1, P2 (C1, the ensemble's groups)
    "k_face_flux_lds<1,double>": (12, 0),
    "k_elem_mfma<1,double,true,false,false>": (12, 0),
    "k_face_flux_dpp<2,double,128>": (12, 0),
    "k_elem_mfma<2,double,true,false,false>": (12, 0),
    "k_face_tile<2,float>": (8, 0),
    "k_elem_mfma<2,float,true,true,false>": (16, 0),
}
LDS_PER_CU = 160 * 1024


def kernel_metadata(tmp_path):
    """{demangled name: dict(vgpr, agpr, spill, scratch, lds)} of every gfx950 kernel in the library"""
    if not LIB.exists():
        pytest.skip("libfvdg.so is not built")
    if not all(t.exists() for t in TOOLS) or shutil.which("c++filt") is None:
        pytest.skip("ROCm LLVM tools or c++filt missing")
    fatbin = tmp_path / "fatbin.bin"
    subprocess.run([str(TOOLS[0]), f"--dump-section=.hip_fatbin={fatbin}", str(LIB), str(tmp_path / "lib.so")],
                   check=True, capture_output=True)
    data = fatbin.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    notes = ""
    for i, (a, b) in enumerate(zip(starts, starts[1:] + [len(data)])):
        bundle, obj = tmp_path / f"b{i}.bin", tmp_path / f"co{i}.o"
        bundle.write_bytes(data[a:b])
        r = subprocess.run([str(TOOLS[1]), "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                            f"--input={bundle}", f"--output={obj}", "--unbundle"], capture_output=True)
        if r.returncode == 0 and obj.exists() and obj.stat().st_size > 0:
            notes += subprocess.run([str(TOOLS[2]), "--notes", str(obj)], check=True, capture_output=True,
                                    text=True).stdout
    rows = []
    for entry in re.split(r"\n\s*- \.agpr_count:", notes)[1:]:
        agpr = int(entry.split("\n", 1)[0].strip())
        f = dict(re.findall(r"^    \.(name|vgpr_count|vgpr_spill_count|private_segment_fixed_size|"
                            r"group_segment_fixed_size|max_flat_workgroup_size):\s+(\S+)", entry, re.M))
        rows.append((f["name"], dict(vgpr=int(f["vgpr_count"]), agpr=agpr, spill=int(f["vgpr_spill_count"]),
                                     scratch=int(f["private_segment_fixed_size"]),
                                     lds=int(f["group_segment_fixed_size"]),
                                     threads=int(f["max_flat_workgroup_size"]))))
    names = subprocess.run(["c++filt"], input="\n".join(n for n, _ in rows), capture_output=True, text=True,
                           check=True).stdout.split("\n")
    return {n.replace("void ", "").split("(")[0].replace(" ", ""): m for n, (_, m) in zip(names, rows)}


def waves_per_simd(m):
    """resident wavefronts per SIMD that the 512-entry register file allows (granule 8)"""
    return min(8, 512 // (-(-(m["vgpr"] + m["agpr"]) // 8) * 8))


def waves_per_cu(m):
    """resident wavefronts per compute unit (four SIMDs): whole blocks, limited by registers and LDS"""
    per_block = -(-m["threads"] // 64)
    blocks = 4 * waves_per_simd(m) // per_block
    if m["lds"]:
        blocks = min(blocks, LDS_PER_CU // m["lds"])
    return blocks * per_block


def test_stage_kernels_within_budget(tmp_path):
    meta = kernel_metadata(tmp_path)
    assert len(meta) > 100, "too few kernels found in the library"
    bad = []
    for name, (waves, scratch) in BUDGET.items():
        assert name in meta, f"{name} not in the library"
        m = meta[name]
        w = waves_per_cu(m)
        print(f"{name}: {m['vgpr']} VGPRs + {m['agpr']} AGPRs, LDS {m['lds']} B per {m['threads']}-thread block "
              f"-> {w} waves per CU; scratch {m['scratch']} B")
        if w < waves or m["scratch"] > scratch:
            bad.append((name, w, waves, m["scratch"], scratch))
    assert not bad, f"(kernel, waves per CU, budget, scratch, budget): {bad}"
