"""Extract the reference's own known-answer data into small JSON fixtures.

Reads era-bellman-cuda's test sources as TEXT (nothing is compiled, imported
or executed) and writes the numeric vectors they hold:

  ntt_lde_kat.json   tests/ntt_test.cu:633-725  (correctness_lde_2 / _4 inputs
                     and expected outputs, raw 8x32-bit limbs)
  msm_bases.json     tests/msm_bases.cuh        (g_bases[k] = (k+1)*G, canonical
                     affine coordinates, 256 points)

Usage: python tests/golden/make_fixtures.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def limbs_to_int(limbs):
    return sum(int(v, 16) << (32 * i) for i, v in enumerate(limbs))


def parse_storage_list(block):
    groups = re.findall(r"\{(0x[0-9a-fA-F]+(?:\s*,\s*0x[0-9a-fA-F]+){7})\}", block)
    return [[x.strip() for x in g.split(",")] for g in groups]


def extract_lde(text, name):
    start = text.index("TEST_F(ntt_test, %s)" % name)
    body = text[start:text.index("correctness_lde(", start)]
    inputs_block = body[body.index("storage inputs[]"):body.index("storage expected_outputs[]")]
    expected_block = body[body.index("storage expected_outputs[]"):]
    log_values = int(re.search(r"log_values_count = (\d+)", body).group(1))
    log_ext = int(re.search(r"log_extension_degree = (\d+)", body).group(1))
    return {
        "log_values_count": log_values,
        "log_extension_degree": log_ext,
        "inputs": [hex(limbs_to_int(l)) for l in parse_storage_list(inputs_block)],
        "expected_outputs": [hex(limbs_to_int(l)) for l in parse_storage_list(expected_block)],
        "source": "tests/ntt_test.cu %s" % name,
    }


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    ntt_text = open(os.path.join(ref, "tests", "ntt_test.cu")).read()
    lde = {name: extract_lde(ntt_text, name) for name in ("correctness_lde_2", "correctness_lde_4")}
    with open(os.path.join(HERE, "ntt_lde_kat.json"), "w") as f:
        json.dump(lde, f, indent=1)

    bases_text = open(os.path.join(ref, "tests", "msm_bases.cuh")).read()
    body = bases_text[bases_text.index("g_bases[]"):]
    coords = parse_storage_list(body)
    assert len(coords) == 512, len(coords)
    points = [[hex(limbs_to_int(coords[2 * i])), hex(limbs_to_int(coords[2 * i + 1]))] for i in range(256)]
    with open(os.path.join(HERE, "msm_bases.json"), "w") as f:
        json.dump({"source": "tests/msm_bases.cuh", "multiple_of_G": "k+1", "points": points}, f, indent=0)


if __name__ == "__main__":
    main()
