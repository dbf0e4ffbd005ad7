"""Bloch unfolding of lead wave vectors (``dpnegf.negf.bloch.Bloch``).

A lead whose principal layer repeats a smaller cell B0 x B1 x B2 times has its
surface Green's function at k assembled from the small cell's at the
B0 B1 B2 points that fold onto k (``dpnegf/negf/bloch.py:25-58``); the
assembly itself is ``lead_property.bloch_self_energy``.
"""
import numpy as np


class Bloch(object):
    def __init__(self, bloch_factor):
        bloch_factor = np.asarray(bloch_factor)
        assert bloch_factor.shape[0] == 3, "kpoint should be a 3D vector"
        self.bloch_factor = bloch_factor

    def unfold_points(self, k):
        """[B2 B1 B0, 3] points ((k_x + i) / B0, (k_y + j) / B1, (k_z + l) / B2),
        the x index running fastest."""
        if isinstance(k, list):
            assert len(k) == 3, "kpoint should be a 3D vector"
        elif isinstance(k, np.ndarray):
            assert k.shape[0] == 3, "kpoint should be a 3D vector"
        else:
            raise ValueError("k should be a list or numpy array")
        B = [int(b) for b in self.bloch_factor]
        grid = np.stack(np.meshgrid(np.arange(B[2]), np.arange(B[1]), np.arange(B[0]),
                                    indexing="ij"), axis=-1).reshape(-1, 3)[:, ::-1]
        return (grid + np.asarray(k, dtype=float)) / np.asarray(B, dtype=float)
