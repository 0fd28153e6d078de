"""ED-MRDT per-timestep particle path on MI355X (gfx950).

Host package mirroring the reference's ``ed_mrdt`` interface (config,
particles, dynamics, photophysics, microscope, pipeline).  The engines run
on the GPU through the C-ABI library built from ``csrc/`` (libedm_hip.so);
there is no CPU fallback — constructing an engine without the library
raises.
"""

__version__ = "1.1.0+mi355x"
