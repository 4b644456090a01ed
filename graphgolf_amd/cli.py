"""Command-line programs with the reference's flags and file formats
(general/main.c:7-95 and :317-497; grid/main.c:3-104 and :586-777), on the
device evaluator.

    python -m graphgolf_amd general -f <edge_file> [-o <output_file>] [-s <random_seed>]
        [-n <num_calculations>] [-w <max_temperature>] [-c <min_temperature>] [-g <groups>]
        [-v <added_centers>] [-B] [-H] [-L] [-M]
    python -m graphgolf_amd grid [-f <edge_file>] [-W <width>] [-H <height>] [-D <degree>]
        -R <length> [-o <output_file>] [-s <random_seed>] [-n <calculations>]
        [-w <max_temperature>] [-c <min_temperature>] [-g <groups>] [-C <cooling_cycle>]
        [-B] [-F <fixed_temperature>] [-Y] [-M]

Edge files are the reference's ("a b" lines; grid "x,y x,y" lines).  A run
follows the reference program's random() stream from the same seed: the same
start-graph moves (skipped with -M, which resumes from a whole graph written
by -o), the same annealing moves, the same output file.  -B and -L choose the
reference's evaluation algorithm.  The graphs these programs evaluate are
undirected and invariant under their rotation, where bfs() and matrix_op give
the same diameter and ASPL (bit for bit), so the device evaluator serves all
three with the same kernels and the flags change the banner only.
-D / -d print the proposed temperatures as the reference does; -N lets
repeated edges stay in the graph (the adjacency is rebuilt after each move).  -O (listed by the
reference's help, absent from its option string) skips the graph checks.
"""
from __future__ import annotations

import math
import os
import sys
import time

import numpy as np

DEFAULT_MAX_TEMP = 100.0      # general/main.c:327, grid/main.c:627
DEFAULT_MIN_TEMP = 0.217147
DEFAULT_NCALCS = 10000


class CliError(Exception):
    """A reference ERROR(): message, exit status 1."""


def _getopt(argv, spec):
    """POSIX getopt over `spec` ("f:o:B..."): returns {flag: value or True}."""
    import getopt
    try:
        opts, rest = getopt.getopt(argv, spec)
    except getopt.GetoptError as exc:
        raise CliError(str(exc)) from None
    if rest:
        raise CliError(f"unexpected argument {rest[0]!r}")
    out = {}
    for k, v in opts:
        out[k[1]] = v if v != "" else True
    return out


def _int(o, k, default, check=None, msg=""):
    if k not in o:
        return default
    try:
        v = int(o[k])
    except ValueError:
        raise CliError(f"-{k} needs an integer") from None
    if check is not None and not check(v):
        raise CliError(msg)
    return v


def _float(o, k, default, check=None, msg=""):
    if k not in o:
        return default
    try:
        v = float(o[k])
    except ValueError:
        raise CliError(f"-{k} needs a number") from None
    if check is not None and not check(v):
        raise CliError(msg)
    return v


def verify_regular(nodes, degree, edges, out):
    """verfy_graph for groups = 1 (general/main.c:212-241): every vertex of
    the degree, no self-loop."""
    print("Verifing a regular graph... ", end="", file=out)
    e = np.asarray(edges).reshape(-1, 2)
    if (2 * len(e)) % degree:
        raise CliError(f"NG. lines or n or d is invalid. lines = {len(e)} nodes = {nodes} "
                       f"degree = {degree}")
    deg = np.bincount(e.ravel(), minlength=nodes)
    bad = np.nonzero(deg != degree)[0]
    if len(bad):
        i = int(bad[0])
        raise CliError(f"NG\nNot regular graph. degree = {degree} n[{i}] = {int(deg[i])}")
    if (e[:, 0] == e[:, 1]).any():
        raise CliError("NG\nThe same node in the edge.")
    print("OK", file=out)


def verify_grid(nodes, degree, edges, height, length, out):
    """verfy_graph (grid/main.c:446-472): every vertex of the degree, every
    edge within the length."""
    print("Verifing a regular graph... ", end="", file=out)
    e = np.asarray(edges).reshape(-1, 2)
    deg = np.bincount(e.ravel(), minlength=nodes)
    bad = np.nonzero(deg != degree)[0]
    if len(bad):
        i = int(bad[0])
        raise CliError(f"NG\nNot regular graph. degree = {degree} n[{i}] = {int(deg[i])}.")
    dist = np.abs(e[:, 0] // height - e[:, 1] // height) + np.abs(e[:, 0] % height - e[:, 1] % height)
    over = np.nonzero(dist > length)[0]
    if len(over):
        i = int(over[0])
        raise CliError(f"Over length in line {i + 1}: length = {length}, distance = {int(dist[i])}")
    print("OK", file=out)


def _outfile_free(path):
    if path and os.path.exists(path):  # general/main.c:444-451, grid/main.c:727-734
        raise CliError(f"Output file {path} exsits. ")


def _detected(r, outfile, out):
    """-D / -d: the proposed temperatures (general/main.c:464-470,
    grid/main.c:749-755).  The reference has created the output file by then
    and leaves it empty."""
    print(f"Proposed max temperature is {(-1.0 * r.max_diff_energy) / math.log(0.5):f}", file=out)
    print(f"Proposed min temperature is {(-2.0) / math.log(0.0001):f}", file=out)
    if outfile:
        open(outfile, "w").close()
    return r


def _results(r, low_diam, low_aspl, elapsed, sep, out):
    print("---", file=out)
    print(f"Diam. k = {r.best_diameter}  ASPL l = {r.best_aspl:f}  Diam. gap = "
          f"{r.best_diameter - low_diam}  ASPL gap = {r.best_aspl - low_aspl:f}", file=out)
    print(f"Steps: {r.steps}{sep} Elapse time: {elapsed:f} sec. (APSP: {r.seconds_eval:f} sec. "
          f"Other: {elapsed - r.seconds_eval:f} sec.)", file=out)


def run_general(argv, evaluator=None, out=sys.stdout, incremental=None):
    """The general program (general/main.c:317-497).  Returns the SaResult
    (edges = the best graph).  `evaluator` (adj, based_nodes, groups,
    added_centers -> (connected, diameter, S)) replaces the device evaluator
    when given.  `incremental` (None: from n >= 4096 with degree <= 32 on the
    device evaluator; groups = 1 without centres) scores proposals with the
    incremental scorer; its results are the same exact integers, so the run
    and the output file are the same."""
    from . import graphs as G
    from . import sa as SA

    o = _getopt(argv, "f:o:s:n:w:c:g:v:BDHLMNOh")
    if "h" in o or "f" not in o:
        raise CliError(__doc__)
    seed = _int(o, "s", 0, lambda v: v >= 0, "-s value >= 0")
    ncalcs = _int(o, "n", DEFAULT_NCALCS, lambda v: v > 0, "-n value >= 1")
    max_t = _float(o, "w", DEFAULT_MAX_TEMP, lambda v: v > 0, "-w value > 0")
    min_t = _float(o, "c", DEFAULT_MIN_TEMP, lambda v: v > 0, "MIN value > 0")
    groups = _int(o, "g", 1, lambda v: v >= 1, "-g value >= 1")
    centers = _int(o, "v", 0, lambda v: v > 0, "-v value > 0")
    hill, halfway, verify = "H" in o, "M" in o, "O" not in o
    if hill and "w" in o:
        raise CliError("Both -H and -w cannot be used.")
    if hill and "c" in o:
        raise CliError("Both -H and -c cannot be used.")
    if hill and "D" in o:
        raise CliError("Both -H and -D cannot be used.")
    if "D" in o and groups != 1:
        raise CliError("When using -D option, -g must be 1.")
    detect, not_simple = "D" in o, "N" in o
    infile, outfile = o["f"], o.get("o")
    if not os.path.exists(infile):
        raise CliError("File not found")
    edges = G.read_edges(infile)
    # general/main.c:361-422
    file_lines = len(edges)
    lines = file_lines if halfway else file_lines * groups
    based_nodes = int(edges.max()) + 1
    if halfway:
        if centers:
            based_nodes -= centers
        if based_nodes % groups:
            raise CliError("based_nodes must be divisible by groups")
        based_nodes //= groups
    nodes = based_nodes * groups
    degree = 2 * lines // nodes
    if nodes <= degree:
        raise CliError(f"n is too small. nodes = {nodes} degree = {degree}")
    source = SA.GENERAL_FULL_FILE if halfway else SA.GENERAL_BASE_FILE
    if not halfway and centers:
        degree += 1
        nodes += centers
        if centers > based_nodes:
            raise CliError(f"-v value ({centers}) is invalid")
        if (degree * nodes) % 2:
            raise CliError(f"Number of degrees or number of nodes is invalid {degree} {nodes}")
        if (nodes - centers) // groups % 2:
            raise CliError("Number of nodes, groups, or added_centers is invalid 1")
        if ((nodes - centers) // groups - degree // groups * centers) % 2:
            raise CliError("Number of nodes, groups, or added_centers is invalid 2")
        source = SA.GENERAL_BASE_CENTRES
    elif halfway and centers:
        nodes += centers
    symmetric = groups > 1 or centers > 0 or detect or not_simple
    if verify and not symmetric:
        verify_regular(nodes, degree, edges, out)
    low_diam, low_aspl = G.moore_bound(nodes, degree)
    rate = 1.0 if hill else (math.pow(min_t / max_t, 1.0 / ncalcs) if max_t != min_t else 1.0)
    _outfile_free(outfile)
    algo = "BFS" if "B" in o else "MATRIX Operation (MEMORY SAVING)" if "L" in o \
        else "MATRIX Opetation"
    print(f"Seed: {seed}", file=out)
    if hill:
        print("Algorithm: Hill climbing Method", file=out)
    else:
        print(f"Algorithm: Simulated Annealing\n   MAX Temperature: {max_t:f}\n"
              f"   MIN Temperature: {min_t:f}\n   Cooling Cycle: 1\n"
              f"   Cooling Rate : {rate:f}", file=out)
    print(f"APSP: {algo}\nNum. of Calulations: {ncalcs}\nInput filename: {infile}\n"
          f"   Vertexes: {nodes}\n   Degree:   {degree}\n   Groups:   {groups}", file=out)
    if centers:
        print(f"   Center points:    {centers}", file=out)
    if outfile:
        print(f"Output filename: {outfile}", file=out)
    print("---", file=out)
    t0 = time.perf_counter()
    if symmetric:
        r = SA.general_program_sym(edges, nodes, degree, groups, centers, source, seed, ncalcs,
                                   max_temp=max_t, min_temp=min_t, hill_climbing=hill,
                                   evaluator=evaluator, not_simple=not_simple,
                                   detect_temp=detect)
    else:
        if incremental is None:
            incremental = evaluator is None and nodes >= 4096 and degree <= 32
        one = None if evaluator is None else (lambda adj: evaluator(adj, nodes, 1, 0))
        r = SA.general_program(edges, nodes, degree, seed, ncalcs, halfway=halfway,
                               max_temp=max_t, min_temp=min_t, hill_climbing=hill,
                               evaluator=one, incremental=incremental)
    if r.already_optimal:
        print("The input file has already optimum solution.", file=out)
        return r
    if detect:
        return _detected(r, outfile, out)
    _results(r, low_diam, low_aspl, time.perf_counter() - t0, ",", out)
    if outfile:
        G.write_edges(outfile, r.edges)
    if verify:
        verify_regular(nodes, degree, r.edges, out)
    return r


def run_grid(argv, evaluator=None, out=sys.stdout):
    """The grid program (grid/main.c:586-777).  Returns the SaResult.
    `evaluator` (adj, groups, based_nodes, rotate_hash -> (connected,
    diameter, S)) replaces the device evaluator when given."""
    from . import graphs as G
    from . import sa as SA

    o = _getopt(argv, "f:W:H:D:R:o:s:n:w:c:g:C:BdF:YMh")
    if "h" in o:
        raise CliError(__doc__)
    width = _int(o, "W", None, lambda v: v > 0, "-W value > 0")
    height = _int(o, "H", None, lambda v: v > 0, "-H value > 0")
    degree = _int(o, "D", None, lambda v: v > 0, "-D value > 0")
    length = _int(o, "R", None, lambda v: v > 0, "-R value > 0")
    seed = _int(o, "s", 0, lambda v: v >= 0, "-s value >= 0")
    ncalcs = _int(o, "n", DEFAULT_NCALCS, lambda v: v > 0, "-n value >= 1")
    groups = _int(o, "g", 1, lambda v: v in (1, 2, 4), "-g value == 1 or 2 or 4")
    cycle = _int(o, "C", 1, lambda v: v > 0, "-C value > 0")
    max_t = _float(o, "w", None, lambda v: v > 0, "-w value > 0")
    min_t = _float(o, "c", None, lambda v: v > 0, "-c value > 0")
    fixed = _float(o, "F", None, lambda v: v > 0, "-F value > 0")
    hill, halfway, infile, outfile = "Y" in o, "M" in o, o.get("f"), o.get("o")
    if length is None:
        raise CliError("Must need -R")
    if hill and max_t is not None:
        raise CliError("Both -Y and -w cannot be used.")
    if hill and min_t is not None:
        raise CliError("Both -Y and -c cannot be used.")
    if hill and "d" in o:
        raise CliError("Both -Y and -d cannot be used.")
    if not infile and None in (width, height, degree):
        raise CliError('Must set -f or "-W and -H and -D"')
    if halfway and not infile:
        raise CliError("Must set both -M and -f")
    detect = "d" in o
    max_t = DEFAULT_MAX_TEMP if max_t is None else max_t
    min_t = DEFAULT_MIN_TEMP if min_t is None else min_t
    if max_t == min_t:
        raise CliError("The same values in -w and -c.")
    if detect:
        ncalcs = 100  # DEFAULT_DETECT_NCALS (grid/common.h:50)
    g = groups
    if infile:  # grid/main.c:631-668
        if not os.path.exists(infile):
            raise CliError("File not found")
        e, bw, bh = G.read_grid_edges(infile)
        lines = len(e) if halfway else len(e) * g
        based_nodes = int(e.max()) + 1
        if halfway:
            based_nodes //= g
            if g == 2:
                bh //= 2
            elif g == 4:
                bw //= 2
                bh //= 2
        width, height = (bw, bh) if g == 1 else (bw, 2 * bh) if g == 2 else (2 * bw, 2 * bh)
        nodes = based_nodes * g
        degree = 2 * lines // nodes
        source = SA.GRID_FULL_FILE if halfway else SA.GRID_BASE_FILE
    else:
        nodes = width * height
        based_nodes = nodes // g
        lines = nodes * degree // 2
        bw = width // 2 if g == 4 else width
        bh = height if g == 1 else height // 2
        e = None
        source = SA.GRID_LATTICE
    if g == 4 and bw != bh:  # grid/main.c:684-695
        raise CliError(f"When g = 4, width({bw}) must be equal to height({bh}).")
    if g == 4 and width % 2 and height % 2:
        raise CliError(f"When g = 4, width({width}) and height({height}) are divisible by 2.")
    if g == 2 and height % 2:
        raise CliError(f"When g = 2, height({height}) os divisible by 2.")
    if nodes % g:
        raise CliError(f"nodes({nodes}) must be divisible by groups({g})")
    if lines % g:
        raise CliError(f"(nodes*degree/2) must be divisible by groups({g})")
    if bw * bh != based_nodes:
        raise CliError(f"Not grid graph (width {bw} x height {bh} != nodes {based_nodes}).")
    low_diam, low_aspl = G.grid_lower_bound(width, height, degree, length)
    rate = 1.0 if hill else math.pow(min_t / max_t, cycle / ncalcs)
    _outfile_free(outfile)
    print(f"Seed     : {seed}\nAPSP     : {'BFS' if 'B' in o else 'MATRIX Opetation'}", file=out)
    if hill:
        print("Algorithm: Hill climbing Method", file=out)
    else:
        print("Algorithm: " + (f"Fixed Temperature Simulated Annealing : {fixed:f}"
                               if fixed else "Simulated Annealing"), file=out)
        print(f"   MAX Temperature: {max_t:f}\n   MIN Temperature: {min_t:f}\n"
              f"   Cooling Cycle: {cycle}\n   Cooling Rate : {rate:f}", file=out)
    if g != 1:
        print(f"   Groups       : {g}", file=out)
    print(f"Num. of Calulations: {ncalcs}", file=out)
    if infile:
        print(f"Input filename: {infile}", file=out)
    print(f"   (w x h, d, r) = ({width} x {height}, {degree}, {length})", file=out)
    if outfile:
        print(f"Output filename: {outfile}", file=out)
    print("---", file=out)
    t0 = time.perf_counter()
    r = SA.grid_program(e, width, height, degree, length, seed, ncalcs, source, groups=g,
                        max_temp=max_t, min_temp=min_t, cooling_cycle=cycle, hill_climbing=hill,
                        fixed_temp=0.0 if hill or fixed is None else fixed, evaluator=evaluator,
                        detect_temp=detect)
    if r.already_optimal:
        print("The input file has already optimum solution.", file=out)
        return r
    if detect:
        return _detected(r, outfile, out)
    _results(r, low_diam, low_aspl, time.perf_counter() - t0, " ", out)
    if outfile:
        G.write_grid_edges(outfile, r.edges, height)
    verify_grid(nodes, degree, r.edges, height, length, out)
    return r


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in ("general", "grid"):
        print(__doc__, file=sys.stderr)
        return 2
    try:
        (run_general if argv[0] == "general" else run_grid)(argv[1:])
    except CliError as exc:
        print(exc, file=sys.stderr)
        return 1
    return 0
