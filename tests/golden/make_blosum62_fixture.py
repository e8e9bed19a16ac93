"""Writes tests/golden/blosum62_ncbistdaa.txt: the 28x28 BLOSUM62 score table (NCBIstdaa order)
as the reference states it in BLAZE_gpu/util.hpp:30-57.  Run once in a container that has the
reference checkout; the fixture is plain data (28 lines of 28 integers)."""
import pathlib
import re
import sys

ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "BLAZE_gpu" / "util.hpp"
text = ref.read_text()
body = text[text.index("int matrix[28][28]"):]
body = body[: body.index("};") + 2]
rows = re.findall(r"\{([-0-9,\s]+)\}", body)
table = [[int(v) for v in r.split(",")] for r in rows]
assert len(table) == 28 and all(len(r) == 28 for r in table), "unexpected table shape"
out = pathlib.Path(__file__).with_name("blosum62_ncbistdaa.txt")
out.write_text("\n".join(" ".join(str(v) for v in r) for r in table) + "\n")
print("wrote", out)
