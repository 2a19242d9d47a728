"""Build-time safety check of this library's own gfx950 assembly (run on the
CPU by __graft_entry__.build()).  Fails if the device code contains any
scalar-memory store, scalar atomic or scalar-cache write-back/discard, and
prints per-kernel register / LDS / spill figures."""
import re
import subprocess
import sys

FORBIDDEN = re.compile(r"^\s*(s_store_|s_buffer_store|s_scratch_store|s_atomic|s_buffer_atomic|"
                       r"s_dcache_wb|s_dcache_discard)", re.M)


def check(src, flags, hipcc="/opt/rocm/bin/hipcc"):
    cmd = [hipcc, "--offload-arch=gfx950", "--offload-device-only", "-S", "-o", "-", src] + flags
    asm = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
    bad = FORBIDDEN.findall(asm)
    if bad:
        raise SystemExit(f"{src}: forbidden scalar-memory write instructions: {sorted(set(bad))}")
    stats = {}
    for m in re.finditer(r"\.name:\s+(\S+)\n(.*?)(?=\n  - \.|\Z)", asm, re.S):
        pass
    for key in ("vgpr_count", "agpr_count", "sgpr_count", "group_segment_fixed_size",
                "private_segment_fixed_size", "vgpr_spill_count"):
        stats[key] = re.findall(rf"\.{key}:\s+(\d+)", asm)
    names = re.findall(r"\.name:\s+(_Z\S+)", asm)
    return names, stats


if __name__ == "__main__":
    flags = sys.argv[2:]
    names, stats = check(sys.argv[1], flags)
    for i, n in enumerate(names):
        row = {k: (v[i] if i < len(v) else "?") for k, v in stats.items()}
        print(n[:60], row)
