"""Differentiable interfaces over the HIP potentials and the batched stepper
(the counterpart of the reference's timemachine/fe/functional.py, in Torch)."""
