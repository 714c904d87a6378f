"""Golden vectors from the reference's own compiled code.

Runs the reference routines built by oracle/Makefile from cerati/p2r-tests'
sources (oracle/_ref/: propagateToR / KalmanUpdate of
src/propagate-tor-test_acc.v3.cpp, the input generator and the full program of
src/propagate-tor-test_pstl.cpp) and stores their outputs as fixtures:

  ref_block_vectors.npz  propagateToR + KalmanUpdate outputs of the reference
                         on one Matriplex block (bsize 32) for
                         - "c0": the reference's own generated C0 inputs
                           (input_track.h smeared, layers 0..7), and
                         - "wide": numpy-seeded inputs with 10 % smears of the
                           acc.v3 base state (acc.v3.cpp:808-819), charge +-1,
                           hits on cylinders (4 layers),
                         run with the reference's kfact (100/3.8).
  ref_gen_c0_head.npz    first 2 Matriplex blocks of tracks and layer-0 hits of
                         the reference generator (pstl.cpp prepareTracks/Hits).
  c0_summary_ref.txt     the summary lines printed by the reference program at
                         10 events x 9600 tracks x 20 layers, bsize 32.

Usage: make -C oracle && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

BS = 32


def wide_block(seed, nlayer):
    """One Matriplex block of benchmark-style inputs (numpy PCG64, stable)."""
    rng = np.random.default_rng(seed)
    base = np.array([-12.806846618652344, -7.723824977874756, 38.13014221191406, 0.23732035065189902,
                     -2.613372802734375, 0.35594117641448975])
    cov = np.array([6.290299552347278e-07, 4.1375109560704004e-08, 7.526661534029699e-07, 2.0973730840978533e-07,
                    1.5431574240665213e-07, 9.626245400795597e-08, -2.804026640189443e-06, 6.219111130687595e-06,
                    2.649119409845118e-07, 0.00253512163402557, -2.419662877381737e-07, 4.3124190760040646e-07,
                    3.1068903991780678e-09, 0.000923913115050627, 0.00040678296006807003, -7.755406890332818e-07,
                    1.68539375883925e-06, 6.676875566525437e-08, 0.0008420574605423793, 7.356584799406111e-05,
                    0.0002306247719158348])
    par = (base[:, None] * (1 + 0.1 * rng.standard_normal((6, BS)))).astype(np.float32)
    cv = (cov[:, None] * (1 + 0.1 * rng.standard_normal((21, BS)))).astype(np.float32)
    q = rng.choice(np.array([-1, 1], np.int32), BS)
    hits = []
    for l in range(nlayer):
        r = 20.0 + 5.0 * l
        ang = np.arctan2(base[1], base[0]) + 0.01 * l
        pos = np.array([r * np.cos(ang), r * np.sin(ang), base[2] + 0.9 * r])
        hp = (pos[:, None] * (1 + 0.1 * rng.standard_normal((3, BS)))).astype(np.float32)
        hc = np.zeros((6, BS), np.float32)
        hc[[0, 2, 5]] = 1e-4
        hits.append((hp, hc))
    return par, cv, q, hits


def run_ref_block(R, par, cv, q, hits):
    outs = []
    for hp, hc in hits:
        e, p = R.propagate_block(cv.ravel(), par.ravel(), q, hp.ravel())
        ke, kp = R.kalman_block(e, p, hc.ravel(), hp.ravel())
        outs.append(np.concatenate([p, e, kp, ke]))
    return np.stack(outs)


def main():
    O, R = oracle.Oracle(), oracle.Ref()
    assert R.bs == BS
    # c0 block: reference generator inputs
    (nevts, ntrks, nlayer, bs), trk, hits = R.pstl_generate()
    blk = trk[:28 * BS]
    c0_hits = []
    for l in range(8):
        hb = hits[l * 9 * BS:(l + 1) * 9 * BS]
        c0_hits.append((hb[:3 * BS].reshape(3, BS), hb[3 * BS:].reshape(6, BS)))
    c0 = run_ref_block(R, blk[:6 * BS].reshape(6, BS), blk[6 * BS:27 * BS].reshape(21, BS),
                       blk[27 * BS:].view(np.int32).copy(), c0_hits)
    wide = run_ref_block(R, *wide_block(20240601, 4))
    np.savez_compressed(os.path.join(HERE, "ref_block_vectors.npz"), c0=c0, wide=wide, kfact=np.float32(R.kfact),
                        wide_seed=np.int64(20240601))
    np.savez_compressed(os.path.join(HERE, "ref_gen_c0_head.npz"), dims=np.array([nevts, ntrks, nlayer, bs]),
                        trk_blocks=trk[:2 * 28 * BS], hits_layer0=hits[:2 * 9 * BS])
    with open(os.path.join(HERE, "c0_summary_ref.txt"), "w") as f:
        f.write("\n".join(R.pstl_c0_summary()) + "\n")
    print("wrote golden vectors to", HERE)


if __name__ == "__main__":
    main()
