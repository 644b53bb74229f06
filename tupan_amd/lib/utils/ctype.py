"""Process-wide precision switch (tupan/lib/utils/ctype.py:13-20).

As in the reference, ``--use_sp`` on the command line selects the float
build (REAL = float32); the default is float64.  This build adds the choice
of force arithmetic for the float64 build:

  ``--use_split``  split-FP32 pair arithmetic with FP64 sums (Acc, AccJerk)
  ``--use_ieee``   FP64 in the reference's operation order

``force_mode`` holds the selection ("fp64", "split32" or "ieee64").
"""
import sys

import numpy as np

use_sp = "--use_sp" in sys.argv
prec = "float32" if use_sp else "float64"

if "--use_split" in sys.argv:
    force_mode = "split32"
elif "--use_ieee" in sys.argv:
    force_mode = "ieee64"
else:
    force_mode = "fp64"

INT = np.dtype(np.int32) if use_sp else np.dtype(np.int64)
UINT = np.dtype(np.uint32) if use_sp else np.dtype(np.uint64)
REAL = np.dtype(np.float32) if use_sp else np.dtype(np.float64)

ctypedict = {"int": INT, "uint": UINT, "real": REAL}
