"""The command-line drop-in on the reference's sample files (written out in
the reference's text formats): same arguments, same final colors as the
hand-derived trace, output file in the "node color" format."""
import os
import subprocess

import pytest

import dyncolor_amd as d

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(d.__file__), "dyncolor_cli")


def _write(tmp_path, sample, colors_key):
    g = tmp_path / "graph.txt"
    g.write_text("".join(f"{a} {b} 1\n" for a, b in sample["edges"]))
    c = tmp_path / "prevColor.txt"
    c.write_text("".join(f"{i} {x}\n" for i, x in enumerate(sample[colors_key])))
    ch = tmp_path / "cE.txt"
    ch.write_text("".join(f"{a} {b} 2 {i}\n" for a, b, i in sample["changes"]))
    return str(g), str(c), str(ch)


@pytest.mark.parametrize("rule,key", [("independent", "independent_hb_colors"), ("ldf", "ldf_hb_colors")])
def test_cli_sample(tmp_path, sample, rule, key):
    g, c, ch = _write(tmp_path, sample, "colors_hb_dir")
    out = tmp_path / "out.txt"
    r = subprocess.run([CLI, g, str(sample["n"]), str(len(sample["edges"])), c, ch,
                        "--rule", rule, "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Total Time for Vertex Color Update" in r.stdout
    assert "totalInsertion:9" in r.stdout
    got = [int(line.split()[1]) for line in out.read_text().splitlines()]
    assert got == sample["expected"][key]["colors"]


def test_cli_rejects_bad_vertex(tmp_path, sample):
    g, c, ch = _write(tmp_path, sample, "colors_hb_dir")
    with open(ch, "a") as f:
        f.write("3 99 1 1\n")
    r = subprocess.run([CLI, g, "10", "19", c, ch], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "out of range" in r.stderr
