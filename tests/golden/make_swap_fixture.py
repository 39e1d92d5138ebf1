"""Extract the K-row dump held in the reference's vim swap file as a JSON fixture.

The reference ships `.test_k.txt.swp`, the swap file of a `test_k.txt` that its
authors printed from Stiffness::calcStiffness with the debug format of
include/stiffness.h:172-177 ("j: %d, K00, K01, K02 / K10, K11, K12 / K20, K21, K22").
It holds particle 0's K row for a 2D hexagonal plate. This script reads the file
as text (it is data, not code) and writes tests/golden/swap_k_row0.json.

Run in the container that holds the reference:
    python tests/golden/make_swap_fixture.py /root/reference/.test_k.txt.swp
"""
import json
import os
import re
import sys

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/.test_k.txt.swp"
raw = open(src, "rb").read()
# printable runs, in file order
lines = [m.decode() for m in re.findall(rb"[ -~]{5,}", raw)]
num = r"(-?\d+\.\d+)"
row_re = re.compile(rf"^{num}, {num}, {num}$")
j_re = re.compile(rf"^j: (\d+), {num}, {num}, {num}$")
blocks = {}
for k, ln in enumerate(lines):
    m = j_re.match(ln)
    if not m:
        continue
    # vim stores the text lines of a block in reverse order: row2, row1, then "j: ..., row0"
    r1 = row_re.match(lines[k - 1])
    r2 = row_re.match(lines[k - 2])
    assert r1 and r2, (ln, lines[k - 1], lines[k - 2])
    j = int(m.group(1))
    blocks[j] = [[float(m.group(2)), float(m.group(3)), float(m.group(4))],
                 [float(r1.group(1)), float(r1.group(2)), float(r1.group(3))],
                 [float(r2.group(1)), float(r2.group(2)), float(r2.group(3))]]
out = {
    "source": "reference .test_k.txt.swp (particle 0 K row, printed with the format of include/stiffness.h:172-177)",
    "blocks": {str(j): blocks[j] for j in sorted(blocks)},
}
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "swap_k_row0.json")
json.dump(out, open(dst, "w"), indent=1)
print("wrote", dst, "with", len(blocks), "blocks:", sorted(blocks))
