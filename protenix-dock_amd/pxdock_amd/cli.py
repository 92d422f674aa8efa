"""The reference's batch tools over ReusableEngine on the GPU:

    python -m pxdock_amd.cli bdock --receptor_file R --ligand_index I [box or --map_dir] ...
    python -m pxdock_amd.cli bdopt --receptor_file R --ligand_index I [box or --map_dir] ...
    python -m pxdock_amd.cli bdscr --receptor_file R --ligand_index I [--output_file scores.csv]

(engine/src/cli/bdock.cpp, bdopt.cpp, bdscr.cpp; map generation alone is
pxdock_amd.bdgen).  Option names and defaults are the reference's: bdock
searches (seed 0, exhaustiveness 256, max_nsteps 200, relax_nsteps 40,
num_modes 8, min_rmsd 0.5, mc_prune_energy_threshold -1), bdopt optimises
poses (max_niters 240), both on maps from --map_dir or built for the box at
--spacing (default 0.175; not above 1e-5: no maps, and bdopt then scores
atom pairs directly), with --penalty (default 1e6) as the out-of-box slope; bdscr
writes pscore and bscore of every pose.  --config reads "name = value"
lines; --nthreads, --num_readers and --num_writers are accepted and unused
(the work runs on the GPU).  Exit status 0 on success, 1 on an option or
run-time error, with the message on stderr.
"""
import argparse
import logging
import sys

from .bdgen import read_config

VERSION_STR = "pxdock-mi355x 0.1"
_BOX = ("center_x", "center_y", "center_z", "size_x", "size_y", "size_z")


def _parser(tool):
    p = argparse.ArgumentParser(prog=tool, add_help=False, allow_abbrev=False)
    g = p.add_argument_group("Inputs")
    g.add_argument("--scoring_function", default="")
    g.add_argument("--receptor_file", default="")
    g.add_argument("--ligand_index", default="")
    g = p.add_argument_group("Outputs")
    if tool == "bdscr":
        g.add_argument("--output_file", default="scores.csv")
    else:
        g.add_argument("--output_dir", default="docked")
    g.add_argument("--log_file", default="-")
    g.add_argument("--verbosity", type=int, default=1)
    if tool != "bdscr":
        g = p.add_argument_group("Precalculation")
        g.add_argument("--map_dir", default="")
        for k in _BOX:
            g.add_argument("--" + k, type=float, default=None)
        g.add_argument("--spacing", type=float, default=0.175)
        g.add_argument("--penalty", type=float, default=1e6)
        g = p.add_argument_group("Algorithm")
        if tool == "bdock":
            g.add_argument("--seed", type=int, default=0)
            g.add_argument("--exhaustiveness", type=int, default=256)
            g.add_argument("--max_nsteps", type=int, default=200)
            g.add_argument("--relax_nsteps", type=int, default=40)
            g.add_argument("--num_modes", type=int, default=8)
            g.add_argument("--min_rmsd", type=float, default=0.5)
            g.add_argument("--mc_prune_energy_threshold", type=float, default=-1.0)
        else:
            g.add_argument("--max_niters", type=int, default=240)
        g = p.add_argument_group("Computing device")
        g.add_argument("--num_readers", type=int, default=1)
        g.add_argument("--num_writers", type=int, default=1)
    g.add_argument("--nthreads", type=int, default=0)
    g = p.add_argument_group("Configuration file (optional)")
    g.add_argument("--config", default=None)
    g = p.add_argument_group("Information")
    g.add_argument("--help", action="store_true")
    g.add_argument("--version", action="store_true")
    return p


def parse_args(tool, argv):
    """Options after the reference tool's checks; ValueError on a bad option."""
    p = _parser(tool)
    a = p.parse_args(argv)
    if a.help or a.version:
        return p, a
    if a.config is not None:
        given = {tok[2:].split("=", 1)[0] for tok in argv if tok.startswith("--")}
        for k, v in read_config(a.config).items():
            if not hasattr(a, k) or k in ("config", "help", "version"):
                raise ValueError(f"unrecognised option '{k}' in {a.config}")
            if k in given:
                continue
            d = p.get_default(k)
            setattr(a, k, v if isinstance(d, str) else (type(d)(v) if d is not None else float(v)))
    if not a.receptor_file:
        raise ValueError("the option '--receptor_file' is required but missing")
    if not a.ligand_index:
        raise ValueError("the option '--ligand_index' is required but missing")
    if a.nthreads < 0:
        raise ValueError("the argument for option '--nthreads' is invalid")
    if not 0 <= a.verbosity <= 2:
        raise ValueError("the argument for option '--verbosity' is invalid")
    if tool == "bdscr":
        if not a.output_file:
            raise ValueError("the argument for option '--output_file' is invalid")
        return p, a
    if not a.output_dir:
        raise ValueError("the argument for option '--output_dir' is invalid")
    needs_box = not a.map_dir and (tool == "bdock" or a.spacing > 1e-5)
    if needs_box:
        for k in _BOX:
            if getattr(a, k) is None:
                raise ValueError(f"the option '--{k}' is required but missing")
        for k in ("size_x", "size_y", "size_z"):
            if not getattr(a, k) > 0:
                raise ValueError(f"the argument for option '--{k}' is invalid")
    if tool == "bdock":
        for k in ("exhaustiveness", "max_nsteps", "num_modes"):
            if getattr(a, k) <= 0:
                raise ValueError(f"the argument for option '--{k}' is invalid")
        if a.relax_nsteps < 0 or a.min_rmsd < 0:
            raise ValueError("the argument for option '--relax_nsteps' / '--min_rmsd' is invalid")
    elif a.max_niters <= 0:
        raise ValueError("the argument for option '--max_niters' is invalid")
    return p, a


def run(tool, a):
    """The tool's engine calls (bdock.cpp:236-262, bdopt.cpp:208-222, bdscr.cpp:116-122)."""
    from .engine import ReusableEngine
    engine = ReusableEngine(a.receptor_file, a.scoring_function, a.nthreads)
    if tool == "bdscr":
        return engine.evaluate(a.ligand_index, a.output_file)
    if a.map_dir:
        engine.load_maps(a.map_dir)
    elif a.spacing > 1e-5 or tool == "bdock":  # CHECK_PARAMETER_POSITIVE (utility.h:24)
        engine.set_box(*(getattr(a, k) for k in _BOX))
        if a.spacing > 1e-5:
            engine.generate_maps(a.spacing)
    if tool == "bdopt":
        return engine.optimize(a.ligand_index, a.output_dir, a.max_niters, a.penalty)
    return engine.search(a.ligand_index, a.output_dir, a.seed, a.exhaustiveness, a.max_nsteps,
                         a.relax_nsteps, a.num_modes, a.min_rmsd, a.mc_prune_energy_threshold,
                         a.penalty)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else list(argv)
    if not argv or argv[0] not in ("bdock", "bdopt", "bdscr"):
        print("usage: python -m pxdock_amd.cli {bdock,bdopt,bdscr} [options]", file=sys.stderr)
        return 1
    tool, argv = argv[0], argv[1:]
    try:
        p, a = parse_args(tool, argv)
    except SystemExit:
        return 1
    except (ValueError, OSError) as e:
        print(e, file=sys.stderr)
        return 1
    if a.version or a.help:
        print(VERSION_STR)
        if a.help:
            p.print_help()
        return 0
    level = (logging.WARNING, logging.INFO, logging.DEBUG)[a.verbosity]
    if a.log_file == "-":
        logging.basicConfig(stream=sys.stdout, level=level, format="%(message)s")
    else:
        logging.basicConfig(filename=a.log_file, level=level, format="%(message)s")
    try:
        run(tool, a)
        return 0
    except Exception as e:  # the tools report any failure and exit 1
        print(e, file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())
