"""Static instruction counts of the four-lane fused pair kernel, per phase.

Compiles csrc/nepx_quad.hip for gfx950 to device assembly (same flags as build.py), picks the
k_fused_quad instance of the given descriptor shape (number of sites per body, hypers), network
precision and activation, and counts the instructions of one iteration of the kernel's main loop
(16 pairs per wave).  Both site loops are unrolled, so the iteration is cut at the `s_setprio 1`
that closes each site trip's block of 4x4x1 matrix products:

  forward   the trips of the forward site loop (site geometry, radial values, harmonics, LDS
            exchange, products); the first trip also holds the iteration's setup (pair loads,
            rotation matrices, pair geometry)
  network   from the last forward block to the first adjoint block: descriptor epilogue, the
            network's 16x16x32 products, adjoint prelude, and the first adjoint trip's radial part
  adjoint   the later adjoint trips (each: the previous trip's gradient terms and sums, this
            trip's geometry and radial part, its c/d products)
  tail      after the last adjoint block: its gradient terms, per-pair sums, clip, stores

Usage: python anisotropic-ml_amd/isa_phases.py [--sites 10] [--hypers 11 6 4] [--prec 1] [--act 0]
"""
import argparse
import collections
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def device_asm(out, source=None):
    sys.path.insert(0, HERE)
    import build
    flags = [f for f in build.FLAGS if f != "-fPIC"]
    cmd = ([build.HIPCC] + flags + build.SOURCE_FLAGS.get("nepx_quad.hip", []) +
           ["--cuda-device-only", "-S", source or os.path.join(build.CSRC, "nepx_quad.hip"), "-o", out])
    subprocess.run(cmd, check=True, capture_output=True)


def kernel_lines(lines, sites, hyp, prec, act):
    tag = f"DescCfgILi{sites}ELi{hyp[0]}ELi{hyp[1]}ELi{hyp[2]}EEE"
    suffix = f"ELi{prec}ELi{act}EEEv"
    start = next(i for i, l in enumerate(lines)
                 if l.startswith("_ZN4nepx12k_fused_quad") and tag in l and suffix in l and ":" in l)
    end = next(i for i in range(start + 1, len(lines)) if lines[i].strip().startswith(".Lfunc_end"))
    return lines[start:end]


def kind(op):
    if op.startswith("v_mfma"):
        return "mfma"
    if op.startswith("v_"):
        return "valu"
    if op.startswith("ds_"):
        return "lds"
    if op.startswith(("global_", "buffer_")):
        return "vmem"
    return "other"


def main_loop(body):
    """Instructions of the main loop's body, in order: from the target of the last backward branch
    to that branch."""
    labels, ins = {}, []
    for l in body[1:]:
        m = re.match(r"^(\.LBB\S+):", l)
        if m:
            labels[m.group(1)] = len(ins)
            continue
        t = l.strip()
        if t and not t.startswith((";", ".")):
            ins.append(t)
    back = [(k, labels[x.split()[-1]]) for k, x in enumerate(ins)
            if x.startswith(("s_branch", "s_cbranch")) and labels.get(x.split()[-1], len(ins)) < k]
    if not back:
        raise SystemExit("no main loop found")
    end, head = max(back, key=lambda b: b[0] - b[1])
    return ins[head:end + 1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sites", type=int, default=10)
    ap.add_argument("--hypers", type=int, nargs=3, default=(11, 6, 4))
    ap.add_argument("--prec", type=int, default=1, help="0 f32, 1 f16x3, 2 bf16, 3 bf16x3")
    ap.add_argument("--act", type=int, default=0, help="0 tanh, 1 relu")
    ap.add_argument("--source", default=None, help="another copy of nepx_quad.hip to count (default: the build's)")
    a = ap.parse_args()
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "quad.s")
        device_asm(out, a.source)
        lines = open(out).read().split("\n")
    body = kernel_lines(lines, a.sites, a.hypers, a.prec, a.act)
    vgpr = next((l.split()[-1] for l in body if ".amdhsa_next_free_vgpr" in l), "?")
    loop = main_loop(body)
    trips = -(-2 * a.sites // 4)
    segs, cur = [], []
    for x in loop:
        cur.append(x)
        if x.split()[0] == "s_setprio" and x.split()[-1] == "1":
            segs.append(cur)
            cur = []
    segs.append(cur)
    if len(segs) != 2 * trips + 1:
        raise SystemExit(f"expected {2 * trips} site-trip blocks, found {len(segs) - 1}")
    names = ["forward"] * trips + ["network"] + ["adjoint"] * (trips - 1) + ["tail"]
    phases = collections.OrderedDict((p, collections.Counter()) for p in ("forward", "network", "adjoint", "tail"))
    for n, seg in zip(names, segs):
        phases[n].update(kind(x.split()[0]) for x in seg)
    print(f"k_fused_quad sites/body={a.sites} hypers={tuple(a.hypers)} prec={a.prec} act={a.act}: "
          f"{vgpr} VGPRs; site trips per iteration = {trips} (unrolled)")
    print("| phase | VALU | MFMA | LDS | VMEM | segments |")
    print("|---|---:|---:|---:|---:|---:|")
    tot = 0
    for p, c in phases.items():
        tot += c["valu"]
        print(f"| {p} | {c['valu']} | {c['mfma']} | {c['lds']} | {c['vmem']} | {names.count(p)} |")
    print(f"| total per 16-pair wave iteration | {tot} | | | | |")


if __name__ == "__main__":
    main()
