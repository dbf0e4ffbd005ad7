"""MI355X implementation of the ``dpnegf.negf`` hot path (same call signatures)."""
