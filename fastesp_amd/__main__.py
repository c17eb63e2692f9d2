"""``python -m fastesp_amd`` -- see fastesp_amd.cli."""
import sys

from .cli import main

sys.exit(main())
