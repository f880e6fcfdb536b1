"""Writes tests/golden/fichera_components.json: the reference's golden for
the DOF numbering of the refined Fichera corner
(unittests/gsBasisComponents_test.cpp:41-79): the number of topology
components per dimension (26 vertices, 51 edges, 33 faces, 7 volumes) and,
for each of the 117 free DOFs of the once-refined degree-1 multibasis
(no Dirichlet sides, glued interfaces), the index of the component it
belongs to in gsBoxTopology::allComponents() order.  Run in the build
container only (it reads the reference tree); the JSON is the committed
fixture."""
import json
import os
import re
import sys

SRC = "/root/reference/unittests/gsBasisComponents_test.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fichera_components.json")


def main():
    text = open(SRC).read()
    counts = re.search(r"numbers_check\s*<<([^;]*);", text).group(1)
    golden = re.search(r"globalIndices_check\s*<<([^;]*);", text, re.S).group(1)
    counts = [int(s) for s in counts.split(",")]
    golden = [int(s) for s in golden.split(",")]
    if len(golden) != 117 or len(counts) != 4:
        sys.exit("golden not found")
    json.dump({"source": "unittests/gsBasisComponents_test.cpp:41-79 (filedata/volumes/fichera.xml)",
               "components_per_dim": counts, "component_of_dof": golden}, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
