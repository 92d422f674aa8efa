"""Map generator: builds the receptor maps of a box on the GPU and saves them
as a map directory (the bdgen tool, engine/src/cli/bdgen.cpp:28-186).

    python -m pxdock_amd.bdgen --receptor_file rec.json \\
        --center_x 1 --center_y 2 --center_z 3 --size_x 22.5 --size_y 22.5 --size_z 22.5 \\
        [--spacing 0.375] [--max_size 0] [--output_dir docked] [--types c3,ca,n]

Options keep bdgen's names, defaults and checks: the six box options and
--receptor_file are required, sizes must be positive, spacing >= 0.01, and
each map size is max(size, max_size) (ReusableEngine::generate_maps,
entrance.cpp:103-121).  --config reads the same options from a file of
"name = value" lines ('#' starts a comment), as boost::program_options'
parse_config_file does; command-line values win.  --scoring_function and
--nthreads are accepted for compatibility: the maps do not depend on the
scoring function and the build runs on the GPU.  --types (not in bdgen,
which always writes all 99 types) restricts the LJ maps to a comma-separated
list of GAFF2 tags or indices; the default is every type.  Exit status 0 on
success, 1 on an option or run-time error, with the message on stderr.
"""
import argparse
import logging
import sys

VERSION_STR = "pxdock-mi355x bdgen 0.1"

_BOX = ("center_x", "center_y", "center_z", "size_x", "size_y", "size_z")


def _parser():
    p = argparse.ArgumentParser(prog="bdgen", add_help=False, allow_abbrev=False,
                                description="Precalculate receptor maps on the GPU.")
    g = p.add_argument_group("Inputs")
    g.add_argument("--scoring_function", default="",
                   help="A Json file defining the scoring function (not used by the maps).")
    g.add_argument("--receptor_file", default="", help="A Json file describing the receptor.")
    g.add_argument("--nthreads", type=int, default=0, help="Accepted; the maps build on the GPU.")
    g = p.add_argument_group("Outputs")
    g.add_argument("--output_dir", default="docked", help="Output folder for the maps.")
    g.add_argument("--log_file", default="-", help='Text file for log messages. "-" is STDOUT.')
    g.add_argument("--verbosity", type=int, default=1, help="0=warning, 1=info, 2=debug")
    g = p.add_argument_group("Precalculation parameters")
    for k in _BOX:
        g.add_argument("--" + k, type=float, default=None)
    g.add_argument("--max_size", type=float, default=0.0,
                   help="map_size_x/y/z = max(max_size, size_x/y/z)")
    g.add_argument("--spacing", type=float, default=0.375, help="Grid spacing >= 0.01 (A).")
    g.add_argument("--types", default="", help="GAFF2 tags or indices to build (default: all).")
    g = p.add_argument_group("Configuration file (optional)")
    g.add_argument("--config", default=None, help="The above options can be put here.")
    g = p.add_argument_group("Information")
    g.add_argument("--help", action="store_true", help="Display usage summary.")
    g.add_argument("--version", action="store_true", help="Display program version.")
    return p


def read_config(path):
    """name = value lines, '#' comments, blank lines ignored."""
    out = {}
    with open(path) as fh:
        for n, line in enumerate(fh, 1):
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            if "=" not in line:
                raise ValueError(f"{path}:{n}: expected 'name = value'")
            k, v = (s.strip() for s in line.split("=", 1))
            out[k] = v
    return out


def type_indices(spec):
    """GAFF2 indices of a comma-separated list of tags or indices ('' = all)."""
    from .lib import type_tags
    TAGS = type_tags()
    if not spec:
        return list(range(len(TAGS)))
    out = []
    for item in spec.split(","):
        item = item.strip()
        if not item:
            continue
        if item.isdigit():
            t = int(item)
            if not 0 <= t < len(TAGS):
                raise ValueError(f"type index {t} out of range")
        elif item in TAGS:
            t = TAGS.index(item)
        else:
            raise ValueError(f"unknown GAFF2 type tag {item!r}")
        out.append(t)
    return sorted(set(out))


def parse_args(argv):
    """Options after bdgen's checks; raises ValueError with bdgen's wording."""
    p = _parser()
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
            cur = getattr(a, k)
            kind = type(p.get_default(k)) if p.get_default(k) is not None else float
            setattr(a, k, kind(v) if cur is None or not isinstance(cur, str) else v)
    if not a.receptor_file:
        raise ValueError("the option '--receptor_file' is required but missing")
    if a.nthreads < 0:
        raise ValueError("the argument for option '--nthreads' is invalid")
    if not a.output_dir:
        raise ValueError("the argument for option '--output_dir' is invalid")
    if not a.log_file:
        raise ValueError("the argument for option '--log_file' is invalid")
    if not 0 <= a.verbosity <= 2:
        raise ValueError("the argument for option '--verbosity' is invalid")
    for k in _BOX:
        if getattr(a, k) is None:
            raise ValueError(f"the option '--{k}' is required but missing")
    for k in ("size_x", "size_y", "size_z"):
        if not getattr(a, k) > 0:
            raise ValueError(f"the argument for option '--{k}' is invalid")
    if not a.spacing - 0.01 > 0:
        raise ValueError("the argument for option '--spacing' is invalid")
    a.type_list = type_indices(a.types)
    return p, a


def generate(a):
    """generate_maps + dump_maps (entrance.cpp:103-131) on the GPU."""
    import torch

    from . import data, lib
    if not torch.cuda.is_available():
        raise RuntimeError("bdgen: no GPU is visible (the maps are built on the device)")
    log = logging.getLogger("bdgen")
    mol = data.parse_receptor(a.receptor_file)
    size = [max(getattr(a, k), a.max_size) for k in ("size_x", "size_y", "size_z")]
    box = lib.decide_grid_dims([a.center_x, a.center_y, a.center_z], size, a.spacing)
    log.info("Generating receptor cache with grid settings:\n"
             "- Front-left-bottom point: [%.3f, %.3f, %.3f]\n"
             "- Sample point count per dimension: [%d, %d, %d]",
             *box.init_xyz, *box.npts)
    R = lib.Receptor(mol)
    G = lib.Grid(R, box, a.type_list)
    torch.cuda.synchronize()
    G.save(a.output_dir)
    log.info("Saved %d maps to %s", len(a.type_list) + 1, a.output_dir)
    return box


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    try:
        p, a = parse_args(argv)
    except (ValueError, OSError, SystemExit) as e:
        if isinstance(e, SystemExit):
            return 1
        print(e, file=sys.stderr)
        _parser().print_help()
        return 1
    if a.version:
        print(VERSION_STR)
        return 0
    if a.help:
        print(VERSION_STR)
        p.print_help()
        return 0
    level = (logging.WARNING, logging.INFO, logging.DEBUG)[a.verbosity]
    if a.log_file == "-":
        logging.basicConfig(stream=sys.stdout, level=level, format="%(message)s")
    else:
        logging.basicConfig(filename=a.log_file, level=level, format="%(message)s")
    try:
        generate(a)
        return 0
    except Exception as e:  # bdgen.cpp:177-183: report and exit 1
        print(e, file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())
