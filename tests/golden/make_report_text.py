"""Copies the reference's committed search report text and a sample of the
number texts of its committed JSON files into fixtures.

Sources (isakruas/fbls, data written by its serde_json pretty printer):
public/2026-11-08/refine/72668830/{search,vetting,pixel,ctoi,download,
search.residuals}.json.  Outputs:

* tests/golden/report_72668830.json - search.json verbatim, the layout
  fixture (sorted keys, indent, number text);
* tests/golden/float_text_72668830.json - distinct number texts with a
  fraction or exponent: every one of the small files and every 37th of the
  residuals file.

Run in a container that mounts the reference at /root/reference.
"""
import json
import os
import re
import shutil

SRC = "/root/reference/public/2026-11-08/refine/72668830"
HERE = os.path.dirname(os.path.abspath(__file__))
shutil.copyfile(os.path.join(SRC, "search.json"), os.path.join(HERE, "report_72668830.json"))

token = re.compile(rb'(?<![\w."])-?\d+(?:\.\d+)?(?:e-?\d+)?(?![\w."])')
seen, out = set(), []
for name, stride in (("search.json", 1), ("vetting.json", 1), ("pixel.json", 1),
                     ("ctoi.json", 1), ("download.json", 1), ("search.residuals.json", 37)):
    text = open(os.path.join(SRC, name), "rb").read()
    floats = [t.decode() for t in token.findall(text) if b"." in t or b"e" in t]
    for t in floats[::stride]:
        if t not in seen:
            seen.add(t)
            out.append(t)
with open(os.path.join(HERE, "float_text_72668830.json"), "w") as f:
    json.dump(out, f, indent=0)
print(len(out))
