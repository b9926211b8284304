"""Shared test geometries: (name, system builder, plan keywords, energy, k-point)."""
from angstromcube_amd import synthetic as S


def cases(small_only=False):
    out = [
        ("periodic8_gamma", lambda: S.small_periodic(), dict(r_trunc=6.0), complex(-0.5, 0.05), (0, 0, 0)),
        ("periodic8_k", lambda: S.small_periodic(), dict(r_trunc=6.0), complex(-0.5, 0.05), (0.25, -1 / 3.0, 0.1)),
        ("isolated6_fd16", lambda: S.make_system(7, (6, 6, 6), natoms=5, numax=2, bc=(0, 0, 0)),
         dict(r_trunc=3.0, nsources=(2, 1, 2), fd_range=(8, 8, 8)), complex(-0.3, 0.02), (0, 0, 0)),
        ("mixed_bc", lambda: S.make_system(8, (6, 5, 7), natoms=5, numax=3, bc=(1, 2, 3)),
         dict(r_trunc=3.0, nsources=(2, 2, 1)), complex(-0.3, 0.1), (0.1, 0, 0)),
    ]
    if not small_only:
        out.append(("wrap16_numax5", lambda: S.make_system(9, (16, 16, 16), natoms=16, numax=5, bc=(1, 1, 1)),
                    dict(r_trunc=6.0, nsources=(2, 1, 1)), complex(-0.3, 0.02), (0, 0, 0)))
    return out
