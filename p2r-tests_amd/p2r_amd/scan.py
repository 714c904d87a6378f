"""Parameter scan of the p2r driver with the reference's harness conventions.

The reference's build.py (cerati/p2r-tests build.py:68-80, :237-270) recompiles one
binary per scan point (the sizes are preprocessor macros), runs it, parses the
`done ntracks=N tot time=T (s)` line (build.py:83) and stores the points with
throughput = N / T in result_<tech>_<backend>_<compiler>[_<suffix>].json:

    {"backend": ..., "compiler": ..., "results": [{<scan point as floats>, "throughput": ...}, ...]}

Here the sizes are runtime flags of lib/p2r_bench, so each point is one run of the
same binary.  Same scan names and defaults, same JSON, same skip/append/overwrite
behaviour; read_throughput.py-style statistics work on the saved driver output.

  python -m p2r_amd.scan --ntrks 8192,9600 --bsize 16,32,64 [-o SUFFIX] [--noH2D] [--noD2H]

Differences: threadsperblockx is the workgroup size of the forward kernel and
must be 64, 128 or 256 (default 256; the reference's GPU default of 32 lanes is
half a CDNA wavefront); nthreads (CPU threads of the TBB/OpenMP builds) is
recorded and has no effect on the GPU driver.
"""
import argparse
import collections
import json
import os
import re
import subprocess
import sys

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_EXE = os.path.join(PKG_ROOT, "lib", "p2r_bench")

# scan parameters with their defaults (build.py:68-80, threadsperblockx adapted)
SCAN_PARAMETERS = [
    ("ntrks", 8192),
    ("nevts", 100),
    ("NITER", 5),
    ("bsize", 32),
    ("nlayer", 20),
    ("nthreads", 1),
    ("num_streams", 1),
    ("threadsperblockx", 256),
]
ScanPoint = collections.namedtuple("ScanPoint", [n for n, _ in SCAN_PARAMETERS])
RESULT_RE = re.compile(r"done ntracks=(?P<ntracks>\d+) tot time=(?P<time>\S+) ")


def throughput(log):
    """Tracks per second from the driver's output (first `done ntracks=... tot time=...` line)."""
    for line in log.split("\n"):
        m = RESULT_RE.search(line)
        if m:
            return int(m.group("ntracks")) / float(m.group("time"))
    raise ValueError("no `done ntracks=... tot time=...` line in the output")


def transfers_mode(no_h2d, no_d2h):
    return {(False, False): "both", (True, False): "d2h", (False, True): "h2d", (True, True): "none"}[(no_h2d, no_d2h)]


def command(exe, point, opts):
    """Driver command line for one scan point."""
    return [exe, "--ntrks", str(point.ntrks), "--nevts", str(point.nevts), "--NITER", str(point.NITER),
            "--bsize", str(point.bsize), "--nlayer", str(point.nlayer), "--num-streams", str(point.num_streams),
            "--threads", str(point.threadsperblockx), "--transfers", transfers_mode(opts.noH2D, opts.noD2H),
            "--input", opts.input, "--semantics", opts.semantics]


def output_name(opts):
    return "result_{}.json".format("_".join(filter(None, ["amd", "hip", "hipcc", opts.output])))


def scan_points(opts):
    import itertools
    return [ScanPoint(*p) for p in itertools.product(*[getattr(opts, n) for n, _ in SCAN_PARAMETERS])]


def run(opts, runner=None):
    """Runs every scan point not already in the result file (unless --append / --overwrite) and writes it.
    `runner(cmd) -> stdout` defaults to a subprocess call; returns the result dict."""
    if runner is None:
        def runner(cmd):
            p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, universal_newlines=True)
            if p.returncode != 0:
                print(" ".join(cmd))
                print(p.stdout)
                raise RuntimeError(f"driver exited with {p.returncode}")
            return p.stdout
    fname = os.path.join(opts.outdir, output_name(opts))
    data = {"backend": "hip", "compiler": "hipcc", "results": []}
    if not opts.overwrite and os.path.exists(fname):
        with open(fname) as f:
            data = json.load(f)
    done = set()
    if not opts.append:
        for r in data["results"]:
            done.add(tuple(r[k] for k in sorted(ScanPoint._fields)))
    for point in scan_points(opts):
        key = tuple(float(getattr(point, k)) for k in sorted(ScanPoint._fields))
        if key in done and not opts.dryRun:
            print("Already found this point in result, skipping:", key)
            continue
        cmd = command(opts.exe, point, opts)
        print("Running {} for {}".format(os.path.basename(opts.exe), point))
        if opts.verbose or opts.dryRun:
            print(" ".join(cmd))
            if opts.dryRun:
                continue
        out = runner(cmd)
        if opts.verbose:
            print(out)
        res = {k: float(getattr(point, k)) for k in ScanPoint._fields}
        res["throughput"] = throughput(out)
        data["results"].append(res)
        print("Throughput {} tracks/second".format(res["throughput"]))
    if not opts.dryRun:
        with open(fname, "w") as f:
            json.dump(data, f, indent=2)
    return data


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description="Scan the MI355X p2r driver over sizes and launch choices")
    for name, default in SCAN_PARAMETERS:
        ap.add_argument("--" + name, type=lambda s: [int(x) for x in s.split(",")], default=[default],
                        help=f"comma separated list (default {default})")
    ap.add_argument("--dryRun", action="store_true", help="print the commands, run nothing")
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("-o", "--output", type=str, default="", help="suffix of the output JSON")
    ap.add_argument("--outdir", type=str, default=".")
    ap.add_argument("--overwrite", action="store_true", help="overwrite the output JSON instead of updating it")
    ap.add_argument("--append", action="store_true", help="append points already present instead of skipping")
    ap.add_argument("--noH2D", action="store_true", help="exclude the host-to-device copy from the timing")
    ap.add_argument("--noD2H", action="store_true", help="exclude the device-to-host copy from the timing")
    ap.add_argument("--input", default="baseline", choices=["baseline", "reference"])
    ap.add_argument("--semantics", default="reference", choices=["reference", "chained"])
    ap.add_argument("--exe", default=DEFAULT_EXE)
    return ap.parse_args(argv)


if __name__ == "__main__":
    run(parse_args())
    sys.exit(0)
