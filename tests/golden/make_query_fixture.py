"""Writes tests/golden/queries_ncbistdaa.txt from the reference's benchmark queries
(BLAZE_queries/*.fa): one line per query, "<name> <letters as NCBIstdaa codes>".
Run once in a container that has the reference checkout; the fixture is plain data."""
import pathlib
import sys

# NCBIstdaa letter order (index = code)
STDAA = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"
CODE = {c: i for i, c in enumerate(STDAA)}

ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "BLAZE_queries"
lines = []
for fa in sorted(ref.glob("*.fa")):
    seq = "".join(l.strip() for l in fa.read_text().splitlines() if l and not l.startswith(">"))
    codes = [CODE.get(c.upper(), CODE["X"]) for c in seq]
    lines.append(fa.stem + " " + " ".join(map(str, codes)))
out = pathlib.Path(__file__).with_name("queries_ncbistdaa.txt")
out.write_text("\n".join(lines) + "\n")
print("wrote", out, len(lines), "queries")
