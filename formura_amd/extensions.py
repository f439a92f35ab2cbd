"""Libraries beyond the generated step libraries (batched launches and the
learned surrogate).  `build_all` compiles them in-tree."""
from __future__ import annotations

from typing import Dict


def build_all() -> Dict[str, str]:
    return {}
