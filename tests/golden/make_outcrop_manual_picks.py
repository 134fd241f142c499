"""Write tests/golden/outcrop_manual_picks_offset50.npz: the observed picks the
outcrop script takes from its manual pick cubes for offset 50
(fieldexamples/outcrop/outcrop.m:139-160: crr_all_manual / crl_all_manual of
outcropDisp.mat, rows fstart:fband, page offset/5), for shots 1, 5, 30 and 60.

Run in a checkout that has the reference tree at /root/reference:
    python tests/golden/make_outcrop_manual_picks.py
The .mat file is read as data (scipy.io.loadmat, arrays only)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from swd_amd import field as F  # noqa: E402

SRC = "/root/reference/fieldexamples/outcrop/outcropDisp.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "outcrop_manual_picks_offset50.npz")
OFFSET = 50
SHOTS = np.array([1, 5, 30, 60])          # 1-based shot numbers (is)


def make(src=SRC):
    prob, _ = F.outcrop(offset=OFFSET)
    crr, crl = F.load_manual_cubes(src)
    cr_r = F.manual_picks(crr, prob["fmin"], prob["fmax"], prob["df"], OFFSET, 5)
    cr_l = F.manual_picks(crl, prob["fmin"], prob["fmax"], prob["df"], OFFSET, 5)
    return dict(cr0_r=cr_r[:, SHOTS - 1], cr0_l=cr_l[:, SHOTS - 1], shots=SHOTS, offset=np.array(OFFSET),
                fmin=np.array(prob["fmin"]), fmax=np.array(prob["fmax"]), df=np.array(prob["df"]))


if __name__ == "__main__":
    np.savez_compressed(OUT, **make())
    print("wrote", OUT)
