"""Molecular-dynamics helpers on the device (timemachine/md of the reference)."""
