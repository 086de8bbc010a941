"""Write small fixtures of the reference's output formats (run once, here).

Reads impurityModel/test/referenceOutput/Ni_NiO_50bath/{XAS.dat, RIXS.bin}
and keeps the first lines of each as data:

* refout_XAS_head.dat: header and the first 5 rows of XAS.dat (text as written);
* refout_RIXS_head.npy: the first 6 rows of the float32 RIXS.bin table
  (51 columns: 50 incoming energies).
"""

import os

import numpy as np

REF = "/root/reference/impurityModel/test/referenceOutput/Ni_NiO_50bath"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    with open(os.path.join(REF, "XAS.dat")) as f:
        head = [next(f) for _ in range(6)]
    with open(os.path.join(HERE, "refout_XAS_head.dat"), "w") as f:
        f.writelines(head)
    t = np.fromfile(os.path.join(REF, "RIXS.bin"), dtype=np.float32).reshape(-1, 51)
    np.save(os.path.join(HERE, "refout_RIXS_head.npy"), t[:6].copy())
    print("rows", t.shape)
