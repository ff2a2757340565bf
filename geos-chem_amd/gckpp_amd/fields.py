"""Per-cell record layouts (met, het) shared with the HIP kernels: parsed from
the generated C header csrc/cell_fields.h so Python and device code cannot
disagree.  MET/HET/HIN map field name -> (offset, length)."""
import re
from pathlib import Path

_H = Path(__file__).resolve().parent.parent / "csrc" / "cell_fields.h"


def _parse(prefix, total_name):
    txt = _H.read_text()
    items = [(m.group(1), int(m.group(2)))
             for m in re.finditer(r"\b%s_(\w+) = (\d+)" % prefix, txt)]
    total = int(re.search(r"\b%s = (\d+)" % total_name, txt).group(1))
    out = {}
    for k, (name, off) in enumerate(items):
        end = items[k + 1][1] if k + 1 < len(items) else total
        out[name] = (off, end - off)
    return out, total


MET, NMET = _parse("MET", "NMET")
HET, NHET = _parse("HET", "NHET")
HIN, NHIN = _parse("HIN", "NHIN")   # fullchem_SetStateHet's model fields
NPHOTOL = int(re.search(r"#define NPHOTOL (\d+)", _H.read_text()).group(1))
