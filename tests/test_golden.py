"""The oracle against its committed results on the 64-strand sphere scene
(tests/golden/config1_frame.npz, written by tests/golden/make_golden.py).

This pins the oracle against unintended change: its positions after one
frame, its Newton and line-search counts, and its objective, gradient and
direction at a perturbed state must come out bit for bit. A deliberate
change of the oracle's arithmetic regenerates the file with the script.
"""
import os

import numpy as np

import golden.make_golden as make_golden

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config1_frame.npz")


def test_oracle_reproduces_golden(oracle_lib):
    want = np.load(FIXTURE)
    got = make_golden.compute(oracle_lib)
    for key in want.files:
        assert np.array_equal(got[key], want[key]), key
