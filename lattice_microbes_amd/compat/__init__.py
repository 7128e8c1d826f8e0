"""Compatibility module ``lm`` (see lm.py in this directory)."""
import os
import sys

DIR = os.path.dirname(os.path.abspath(__file__))


def install() -> None:
    """Make ``import lm`` resolve to this directory's lm.py."""
    if DIR not in sys.path:
        sys.path.insert(0, DIR)
