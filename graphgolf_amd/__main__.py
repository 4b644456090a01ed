"""python -m graphgolf_amd {general,grid} ... (graphgolf_amd/cli.py)."""
import sys

from .cli import main

sys.exit(main())
