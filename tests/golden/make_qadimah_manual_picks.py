"""Write tests/golden/qadimah_manual_picks_offset39.npz: the observed picks the
Qadimah-fault script takes from its manual pick cubes for offset 39
(fieldexamples/QadimahFault/QademahSWD3.m:73,149-153: crObsManualSR /
crObsManualSL of manualDispfile3of3.mat, rows fstart:fband, page offset/3,
then smooth2a(., 2, 1) over the whole page, here the oracle's FP64
restatement of core/smooth2a.m), for shots 3, 60 and 118.

Run in a checkout that has the reference tree at /root/reference:
    python tests/golden/make_qadimah_manual_picks.py
The .mat file is read as data (scipy.io.loadmat, arrays only)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import swd_oracle as O  # noqa: E402
from swd_amd import field as F  # noqa: E402

SRC = "/root/reference/fieldexamples/QadimahFault/manualDispfile3of3.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qadimah_manual_picks_offset39.npz")
OFFSET = 39
SHOTS = np.array([3, 60, 118])            # 1-based shot numbers (is)


def make(src=SRC):
    prob, _ = F.qadimah(offset=OFFSET)
    crr, crl = F.load_manual_cubes(src, names=("crObsManualSR", "crObsManualSL"))
    cr_r = O.smooth2a(F.manual_picks(crr, prob["fmin"], prob["fmax"], prob["df"], OFFSET, 3), 2, 1)
    cr_l = O.smooth2a(F.manual_picks(crl, prob["fmin"], prob["fmax"], prob["df"], OFFSET, 3), 2, 1)
    return dict(cr0_r=cr_r[:, SHOTS - 1], cr0_l=cr_l[:, SHOTS - 1], shots=SHOTS, offset=np.array(OFFSET),
                fmin=np.array(prob["fmin"]), fmax=np.array(prob["fmax"]), df=np.array(prob["df"]))


if __name__ == "__main__":
    np.savez_compressed(OUT, **make())
    print("wrote", OUT)
