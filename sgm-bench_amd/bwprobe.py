"""Run the streaming-copy bandwidth probe: python sgm-bench_amd/bwprobe.py [--mib N] [--reps K]."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from sgm_amd.bwprobe import main  # noqa: E402

if __name__ == "__main__":
    main()
