"""``python -m wadi2d_amd run job.json --out DIR``: the run protocol's command line (wadi2d_amd.cli)."""

import sys

from wadi2d_amd.cli import main

sys.exit(main())
