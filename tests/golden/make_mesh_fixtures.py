"""Write the mesh-reader fixtures tests/golden/meshes/<case>.{cns,geo,json}
with the reference mesh generator.

The generator is built from the reference's own sources into oracle/_ref/gen
by `make -C oracle ref` (Generate/ and the Library/ modules it uses), and is
run here, in this container only, on three inputs:
  1solid            Generate/Library_cases/1solid.d as it is (240 cells)
  cavity_small      Generate/Library_cases/cavity.d with its block at 7x7x7
                    nodes instead of 11x11x11 (two refinement levels, cells
                    with up to 12 neighbours)
  channel_periodic  the domain of Test_cases/Problem1/problem1.sh with 9
                    nodes in x (periodic in y and z)
  2mat_small        Generate/Library_cases/2mat.d with its block at 13x9x5
                    nodes instead of 61x31x15 and its second-material box
                    at cells 3..6 x 3..6 x 1..3 (periodic in y)
with the answers problem1.sh gives the generator.  The .json file holds the
grid statistics the generator prints (cells, sides, boundary cells, matrix
entries, the neighbour histogram, periodic links), which the reader tests
use as independent pins.
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
REF = os.environ.get("TFLOWS_REFERENCE", "/root/reference")
GEN = os.path.join(REPO, "oracle", "_ref", "gen")
OUT = os.path.join(HERE, "meshes")
ANSWERS = "gen\n0\n0\nx\nyz\n0\n0\n0\n1 1 4\nC\nBC\nskip\n\n"     # Test_cases/Problem1/problem1.sh:59-72


def domains():
    lib = os.path.join(REF, "Generate", "Library_cases")
    d = {"1solid": open(os.path.join(lib, "1solid.d")).read()}
    cav = open(os.path.join(lib, "cavity.d")).read()
    small = cav.replace("  1   11  11  11 \n", "  1   7  7  7 \n")
    assert small != cav
    d["cavity_small"] = small
    two = open(os.path.join(lib, "2mat.d")).read()
    small2 = two.replace("   1   61  31  15 \n", "   1   13  9  5 \n").replace(
        "    3    11 11  1  20  20  7\n", "    3    3 3  1  6  6  3\n")
    assert small2.count("13  9  5") == 1 and small2.count("3 3  1  6  6  3") == 1
    d["2mat_small"] = small2
    sh = open(os.path.join(REF, "Test_cases", "Problem1", "problem1.sh")).read()
    a = sh.index("cat << EOF > gen.d") + len("cat << EOF > gen.d\n")
    d["channel_periodic"] = sh[a:sh.index("EOF", a)].replace('$(echo "2^$i+1" | bc -l)', "9")
    return d


def stats(text):
    num = lambda pat: int(re.search(pat + r"\s*:\s*(\d+)", text).group(1))   # noqa: E731
    out = {"nc": num(r"number of cells"), "ns": num(r"number of sides"), "nbc": num(r"number of boundary cells"),
           "nnz": num(r"number of non zero matrix entries"),
           "neighbours": {int(k): int(v) for k, v in re.findall(r"cells with\s+(\d+)\s+neighbours:\s+(\d+)", text)}}
    m = re.search(r"Periodic links\s*:\s*(\d+)", text)
    out["periodic_links"] = int(m.group(1)) if m else 0
    return out


def main():
    if not os.path.exists(GEN):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    os.makedirs(OUT, exist_ok=True)
    for name, text in domains().items():
        with tempfile.TemporaryDirectory() as tmp:
            open(os.path.join(tmp, "gen.d"), "w").write(text)
            r = subprocess.run([GEN], input=ANSWERS, capture_output=True, text=True, cwd=tmp, timeout=120)
            if r.returncode != 0:
                sys.exit(f"{name}: generator failed\n{r.stdout[-2000:]}{r.stderr[-2000:]}")
            for ext in ("cns", "geo"):
                shutil.copy(os.path.join(tmp, "gen." + ext), os.path.join(OUT, f"{name}.{ext}"))
            st = stats(r.stdout)
            json.dump(st, open(os.path.join(OUT, f"{name}.json"), "w"), indent=1, sort_keys=True)
            print(name, st)


if __name__ == "__main__":
    main()
