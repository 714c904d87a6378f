"""The build.py-style scan harness (p2r_amd.scan): command lines per scan point, the result JSON layout
(build.py:237-270), skip / append / overwrite of existing points, and the throughput parse of the driver's
`done ntracks=N tot time=T (s)` line.  The driver is replaced by a function returning its output text."""
import json

from p2r_amd import scan


def _opts(tmp_path, *extra):
    return scan.parse_args(["--outdir", str(tmp_path), "--ntrks", "9600,19200", "--bsize", "32,64", *extra])


def _fake_runner(calls):
    def run(cmd):
        calls.append(cmd)
        ntrks = int(cmd[cmd.index("--ntrks") + 1])
        nevts = int(cmd[cmd.index("--nevts") + 1])
        niter = int(cmd[cmd.index("--NITER") + 1])
        n = ntrks * nevts * niter
        return (f"NITER={niter}\nsetup time time=0.1 (s)\n"
                f"done ntracks={n} tot time={n * 1e-9:f} (s) time/trk={1e-9:e} (s)\nformatted ...\n")
    return run


def test_throughput_parse_matches_the_driver_line():
    line = "done ntracks=%i tot time=%f (s) time/trk=%e (s)" % (4800000, 0.25, 0.25 / 4800000)
    assert scan.throughput("x\n" + line + "\n") == 4800000 / 0.25


def test_command_line_of_a_point(tmp_path):
    o = _opts(tmp_path, "--noD2H", "--threadsperblockx", "128")
    p = scan.scan_points(o)[0]
    cmd = scan.command("/x/p2r_bench", p, o)
    assert cmd[0] == "/x/p2r_bench"
    for flag, val in (("--ntrks", "9600"), ("--bsize", "32"), ("--threads", "128"), ("--transfers", "h2d"),
                      ("--num-streams", "1"), ("--NITER", "5")):
        assert cmd[cmd.index(flag) + 1] == val
    assert len(scan.scan_points(o)) == 4


def test_results_json_and_skip_append_overwrite(tmp_path):
    calls = []
    data = scan.run(_opts(tmp_path), runner=_fake_runner(calls))
    f = tmp_path / "result_amd_hip_hipcc.json"
    saved = json.loads(f.read_text())
    assert saved == data and saved["backend"] == "hip" and saved["compiler"] == "hipcc"
    assert len(saved["results"]) == 4 and len(calls) == 4
    r = saved["results"][0]
    assert set(r) == set(scan.ScanPoint._fields) | {"throughput"} and isinstance(r["ntrks"], float)
    assert abs(r["throughput"] - 1e9) / 1e9 < 1e-3
    scan.run(_opts(tmp_path), runner=_fake_runner(calls))  # every point present: nothing runs
    assert len(calls) == 4
    scan.run(_opts(tmp_path, "--append"), runner=_fake_runner(calls))
    assert len(calls) == 8 and len(json.loads(f.read_text())["results"]) == 8
    scan.run(_opts(tmp_path, "--overwrite"), runner=_fake_runner(calls))
    assert len(json.loads(f.read_text())["results"]) == 4


def test_dry_run_runs_nothing(tmp_path):
    calls = []
    scan.run(_opts(tmp_path, "--dryRun", "-o", "x"), runner=_fake_runner(calls))
    assert calls == [] and not (tmp_path / "result_amd_hip_hipcc_x.json").exists()
