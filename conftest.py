import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "ed-lgt_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))
