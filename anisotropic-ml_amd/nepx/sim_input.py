"""Command-line driver with the reference flags (nep_simulator/sim_input.py:18-61).

python -m nepx.sim_input --init init.gsd --out run --ts 20000 --dump 1000 --kT 0.7 --dt 0.005 \\
    --model_path model.pth --gpu_id 0 --hypers_nep 10 8 4 4.5 --shape th_v1
Extra flags: --act (network activation, default relu as the reference network),
--nn_width/--nn_depth (default 150/3 as Sim.py:60), --fd_compat (reference FD forces).
"""
import argparse
import os
import sys


def main(argv=None):
    parser = argparse.ArgumentParser(description="Runs NN-assisted MD")
    g = parser.add_argument_group("mandatory arguments")
    g.add_argument("--init", metavar="<dat>", type=str, dest="init_gsd", required=True, help=".gsd files ")
    g.add_argument("--out", metavar="<dat>", type=str, dest="output", required=True, help=".gsd files ")
    g.add_argument("--ts", metavar="<int>", type=int, dest="timesteps", required=True, help="length of sim ")
    g.add_argument("--dump", metavar="<int>", type=int, dest="dump_period", required=True, help="dump frequency ")
    g.add_argument("--kT", metavar="<float>", type=float, dest="kT", required=True, help="temperature ")
    g.add_argument("--dt", metavar="<float>", type=float, dest="dt", required=True, help="dt ")
    g.add_argument("--hypers_nep", metavar="<float>", type=float, dest="hypers_nep", required=True, nargs="+")
    g.add_argument("--model_path", metavar="<dat>", type=str, dest="model_path", required=True)
    g.add_argument("--gpu_id", metavar="<int>", type=int, dest="gpu_id", required=True)
    g.add_argument("--shape", metavar="<str>", type=str, dest="shape", required=True)
    parser.add_argument("--act", default="relu", choices=["relu", "tanh"])
    parser.add_argument("--nn_width", type=int, default=150)
    parser.add_argument("--nn_depth", type=int, default=3)
    parser.add_argument("--fd_compat", action="store_true")
    args = parser.parse_args(argv)

    from .NVT import NVT

    os.makedirs("./out/", exist_ok=True)
    gsd_filename = "./out/" + args.output + ".gsd"
    sim = NVT()
    sim.nn_width, sim.nn_depth, sim.activation = args.nn_width, args.nn_depth, args.act
    sim.fd_compat = args.fd_compat
    sim.setDevice(args.gpu_id)
    sim.initializeSystem(args.init_gsd)
    sim.setNepDescriptors(args.hypers_nep)
    sim.setEvaluator(args.model_path)
    sim.setShape(args.shape)
    sim.setNeighborList()
    sim.setGsdDump(args.dump_period, gsd_filename)
    sim.setkT(args.kT)
    sim.setdt(args.dt)
    sim.run(args.timesteps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
