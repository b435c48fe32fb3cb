"""Extract the known-answer values held by the reference's fixture output/jax_ref/output_lut_1.npz.

The fixture is the hit list written by optimize/simulate.py:170-179 (``--out_np``) for one batch
of module0 events.  Its hit ticks are partly fractional (an older self-trigger model), so it pins
only the functions that act per hit: ``digitize`` (fee_jax.py:55-71) through ``adc2charge``
(losses_jax.py:380-383, stored as ``Q``), ``get_pixel_coordinates`` (detsim_jax.py:296-306,
stored as ``x``/``y``) and ``get_hit_z`` (detsim_jax.py:308-319, stored as ``z`` beside
``ticks``).  Only the distinct values are kept; the file is data (loaded with
allow_pickle=False), not code.

Usage: python tests/golden/make_known_answers.py [path to output_lut_1.npz]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/output/jax_ref/output_lut_1.npz"

d = np.load(SRC, allow_pickle=False)
zt = np.unique(np.stack([d["z"], d["ticks"]], axis=1), axis=0)
aq = np.unique(np.stack([d["adcs"], d["Q"]], axis=1), axis=0)
np.savez_compressed(os.path.join(HERE, "output_lut_1_known.npz"),
                    x=np.unique(d["x"]), y=np.unique(d["y"]), z=zt[:, 0], ticks=zt[:, 1],
                    adcs=aq[:, 0], Q=aq[:, 1], event_id=np.unique(d["event_id"]),
                    n_hits=np.array([len(d["adcs"])]))
print("wrote", os.path.join(HERE, "output_lut_1_known.npz"), "from", len(d["adcs"]), "hits")
