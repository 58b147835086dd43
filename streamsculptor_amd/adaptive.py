"""Adaptive Dormand-Prince 5(4) integration (placeholder until the kernel lands)."""


def integrate_orbit(*args, **kwargs):
    raise NotImplementedError("adaptive Dopri5 is not built yet; use solver='leapfrog', dt0=...")


def integrate_to(*args, **kwargs):
    raise NotImplementedError("adaptive Dopri5 is not built yet; use solver='leapfrog', dt0=...")
