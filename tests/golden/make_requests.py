"""Write the reference's example research requests as JSON fixtures.

Reads examples/research_{slump,coupled_patch,hydrostatic_water}.yaml of the
reference checkout (data: wire requests, loaded with yaml.safe_load) and writes
requests_<name>.json beside this script.  Run from the repository root:
    python tests/golden/make_requests.py /root/reference
"""
import json
import os
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ("research_slump", "research_coupled_patch", "research_hydrostatic_water")


def main(reference):
    for name in NAMES:
        with open(os.path.join(reference, "examples", name + ".yaml")) as f:
            request = yaml.safe_load(f)
        with open(os.path.join(HERE, f"request_{name}.json"), "w") as f:
            json.dump(request, f, indent=1, sort_keys=True)
            f.write("\n")
        print("wrote", name)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
