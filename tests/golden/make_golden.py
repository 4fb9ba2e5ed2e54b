"""Extracts golden vectors from the reference's shipped output files.

Run in a container where the reference is mounted at /root/reference:

    python tests/golden/make_golden.py

The reference ships outputs of one forward run (CUDA_week2/io/...): the
receiver traces io/receiver/rec_signal.csv (3 receivers x 334 steps) and vz
snapshots io/snap_data/vz_snap{27,39..66}.csv (613 x 623, every 5 steps),
and seven identical Ricker source traces io/receiver/src_signal_1_*.csv
(7 x 2500).  These are data, written by the reference with the default
6-significant-digit ostream format.  This script stores them (or a fixed
subsample of the large snapshots) as compressed numpy archives, so the tests
do not need the reference at run time.

The run configuration they come from is CUDA_week2/test_configuration.ini
with time = 0.2 s (nt = 334), full_waveform_inversion = 0, snap_interval = 5
writing vz, the homogeneous model of fd_medium.cu:61-75 and the Taylor
coefficient hc[1] = 1.0 (max_relative_error = 0).  That reconstruction is
recorded in tests/golden/ref_run.json; it was identified from the data
itself (light-cone onsets, the 1.001^n amplitude ratio of the Holberg
coefficient, and the (lambda+2mu) ratio of the precursors inside the disc).
"""
import json
import os

import numpy as np

REF = "/root/reference/CUDA_week2/io"
OUT = os.path.dirname(os.path.abspath(__file__))


def load_csv(path):
    rows = []
    with open(path) as f:
        for line in f:
            vals = [v for v in (t.strip() for t in line.strip().split(",")) if v]
            rows.append([float(v) for v in vals])
    return rows


def main():
    rec = np.array(load_csv(os.path.join(REF, "receiver", "rec_signal.csv")))
    assert rec.shape == (3, 334)

    src_rows = load_csv(os.path.join(REF, "receiver", "src_signal_1_1.csv"))
    src = np.array(src_rows)
    assert src.shape == (7, 2500)
    assert np.all(src == src[0])

    # Snapshots: keep every 4th row and column plus the full CPML columns
    # (0..11 and 611..622), which are the entries that pin the x strips.
    snaps = {}
    rows = np.arange(0, 613, 4)
    cols = np.unique(np.concatenate([np.arange(0, 623, 4), np.arange(0, 12), np.arange(611, 623)]))
    for k in (39, 52, 66):
        s = np.array(load_csv(os.path.join(REF, "snap_data", f"vz_snap{k}.csv")))
        assert s.shape == (613, 623)
        snaps[f"snap{k}"] = s[np.ix_(rows, cols)]
        snaps[f"snap{k}_strips"] = s[:, np.r_[0:12, 611:623]]
    np.savez_compressed(os.path.join(OUT, "ref_forward_run.npz"), rec_signal=rec,
                        src_signal_row=src[0], snap_rows=rows, snap_cols=cols, **snaps)

    run = {
        "source": "CUDA_week2/test_configuration.ini with time=0.2, full_waveform_inversion=0",
        "nx": 601, "nz": 601, "dx": 4.0, "dz": 4.0, "dt": 0.0006, "nt": 334,
        "snap_interval": 5, "snapshot_field": "vz", "free_surface": 1,
        "scalar_rho": 7100.0, "scalar_vp": 2955.0, "scalar_vs": 2362.0,
        "model": "homogeneous (fd_medium.cu:61-75) on the padded grid",
        "npml": 10, "fpml": 17.0, "dh": 4.0, "npower": 2.0, "k_max_PML": 1.0,
        "rcoef": 0.001, "damp_v_PML": 3300.0, "fdorder": 2, "max_relative_error": 0,
        "source_x": 300, "source_z": 2, "source_amp": 1.0, "source_fc": 17.0,
        "receiver_x": [100, 300, 400], "receiver_z": [0, 0, 0],
        "src_signal_file": {"nt": 2500, "dt": 0.0002, "amp": 1000.0, "fc": 15.0, "ts": 0.0},
    }
    with open(os.path.join(OUT, "ref_run.json"), "w") as f:
        json.dump(run, f, indent=1)
    print("wrote", os.path.join(OUT, "ref_forward_run.npz"))


if __name__ == "__main__":
    main()
