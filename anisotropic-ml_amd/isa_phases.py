"""Static instruction counts of the four-lane fused pair kernel, per phase.

Compiles csrc/nepx_quad.hip for gfx950 to device assembly (same flags as build.py), picks the
k_fused_quad instance of the given descriptor shape (number of sites per body, hypers), network
precision and activation, cuts it into basic blocks and groups them by their place in the
kernel's main loop (one iteration = 16 pairs per wave):

  setup     blocks from the loop head to the forward site loop (pair loads, rotation matrices,
            pair geometry)
  forward   the forward site loop (one site per lane per trip; trips = sites per lane)
  network   the blocks between the two site loops that issue the network's 16x16x32 MFMAs (with
            whatever the compiler schedules into the same basic block, e.g. the descriptor epilogue)
  between   the other blocks between the site loops (adjoint prelude, descriptor epilogue if apart)
  adjoint   the adjoint site loop (same trip count)
  tail      the rest of the iteration (per-pair sums, clip, stores)

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


def blocks_of(body):
    blocks, cur = [], ["entry", []]
    for l in body[1:]:
        m = re.match(r"^(\.LBB\S+):", l)
        if m:
            blocks.append(cur)
            cur = [m.group(1), []]
            continue
        t = l.strip()
        if not t or t.startswith(";") or t.startswith("."):
            continue
        cur[1].append(t)
        if t.startswith(("s_cbranch", "s_branch")):  # a branch ends the block; the fall-through starts one
            blocks.append(cur)
            cur = [cur[0] + "+", []]
    blocks.append(cur)
    return [b for b in blocks if b[1]]


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
    blocks = blocks_of(body)
    # the site loops: blocks that branch back to their own label and issue the 4x4 MFMAs
    loops = [k for k, (lab, ins) in enumerate(blocks)
             if ins[-1].startswith("s_cbranch") and ins[-1].split()[-1] == lab and
             any(x.startswith("v_mfma_f32_4x4x1") for x in ins)]
    if len(loops) < 2:
        raise SystemExit(f"expected two site loops, found {len(loops)}")
    f_loop, a_loop = loops[0], loops[1]
    index = {lab: k for k, (lab, _) in enumerate(blocks)}
    # the iteration's loop head: target of the branches after the adjoint loop that jump back above
    # the forward loop (the last of them ends the tail); the head and the blocks up to the forward
    # loop are the setup
    back = [(k, index[ins[-1].split()[-1]]) for k, (lab, ins) in enumerate(blocks)
            if k > a_loop and ins[-1].startswith(("s_branch", "s_cbranch")) and
            index.get(ins[-1].split()[-1], len(blocks)) < f_loop]
    if not back:
        raise SystemExit("no branch back to the loop head")
    tail_end, head = back[-1]
    phases = collections.OrderedDict((p, collections.Counter()) for p in
                                     ("setup", "forward", "network", "between", "adjoint", "tail"))
    for k, (lab, ins) in enumerate(blocks):
        c = collections.Counter(kind(x.split()[0]) for x in ins)
        if k == f_loop:
            phases["forward"].update(c)
        elif k == a_loop:
            phases["adjoint"].update(c)
        elif head <= k < f_loop:
            phases["setup"].update(c)
        elif f_loop < k < a_loop:
            net = any(x.startswith("v_mfma_f32_16x16x32") for x in ins)
            phases["network" if net else "between"].update(c)
        elif a_loop < k <= tail_end:
            phases["tail"].update(c)
        elif k > tail_end and not ins[-1].startswith("s_endpgm"):
            phases["between"].update(c)  # out-of-line blocks of the region between the loops
    trips = -(-2 * a.sites // 4)
    print(f"k_fused_quad sites/body={a.sites} hypers={tuple(a.hypers)} prec={a.prec} act={a.act}: "
          f"{vgpr} VGPRs; site-loop trips per iteration = {trips}")
    print("| phase | VALU | MFMA | LDS | VMEM | per iteration (VALU) |")
    print("|---|---:|---:|---:|---:|---:|")
    tot = 0
    for p, c in phases.items():
        mult = trips if p in ("forward", "adjoint") else 1
        tot += c["valu"] * mult
        print(f"| {p}{' (per trip)' if mult > 1 else ''} | {c['valu']} | {c['mfma']} | {c['lds']} | {c['vmem']} | "
              f"{c['valu'] * mult} |")
    print(f"| total per 16-pair wave iteration | | | | | {tot} |")


if __name__ == "__main__":
    main()
