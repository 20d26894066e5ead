"""mdforce: MI355X-native P3M Yukawa force engine.

Drop-in modules with the reference's names and signatures:
    yukawa_p3m_ord6.force_pot            yukawa_pp.particle_particle
    yukawa_pm_ord6.particle_mesh_fft_r   yukawa_gf_opt.gf_opt
    velocity_verlet_p3m_ord6.update      thermostat_p3m_ord6.vscale
    initialize_pos_vel.initial
Device API: mdforce.engine.P3M.  All compute runs in libmdforce.so (HIP,
gfx950); there is no CPU fallback.
"""
from ._lib import MdforceError, lib  # noqa: F401

__all__ = ["MdforceError", "lib"]
