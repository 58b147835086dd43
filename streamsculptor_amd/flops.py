"""Algorithmic FP64 operation counts per particle-step, frozen before the first measurement.

Convention: +, -, x, /, sqrt and log1p count 1 FLOP each and a fused multiply-add counts 2;
counts are per formula as written (no sharing between components).  These are the figures the
roofline fractions are computed from; hardware instruction counts (rocprof) exceed them because
FP64 sqrt, division, rsqrt and log1p expand to multi-instruction sequences on CDNA.

  host gradient   MN 20 + Hernquist 16 + NFW 21                      = 57
  KDK update      2 half kicks (2 x 6) + drift (6)                    = 18
  Plummer source  difference 3 + squared radius 6 + rsqrt 2 + cube 3 + accumulate 6 - 2
                  (the squared-radius add of r_s^2 fuses)             = 18 per source
  host Hessian    (response and adjoint forms)                        = 96
  response vector per subhalo: unit-mass Plummer 16 + H dx 15 + source add 3 + KDK 18 = 52
  Dopri5 step (accepted or rejected), per state component: stage combinations of 1..5
                  terms (2m + 1 each: 35) + 5th-order solution 11 + error estimate 13 +
                  scaled error 3 = 62, x 6 components = 372; RMS norm 14 + controller 10
                  = 396, plus 6 right-hand sides of (host gradient + 18 per active source
                  + 6 for [dir v, -dir grad])
  adaptive responses, per particle-step (accepted or rejected): the Dopri5 step of the base
                  row (396) with 6 right-hand sides of host gradient + Hessian (57 + 96 + 6),
                  and per 6-component response vector the same per-component step work
                  (62 x 6 = 372) plus 6 right-hand sides of unit source 16 + H dx 15 +
                  source add 3 = 34: 1,350 + 576 per vector
"""
HOST_GRAD = 57
KDK = 18
PLUMMER_SOURCE = 18
HOST_HESSIAN = 96
RESPONSE_PER_SUBHALO = 52
HERNQUIST_GRAD = 16
DOPRI5_STEP = 396
RESPONSE_RHS = 34


def plain(n_sources: int = 0) -> int:
    """Leapfrog particle-step in host + n Plummer sources (plain workload: 75; 512 clumps:
    9,291; 64 subhalos: 1,227)."""
    return KDK + HOST_GRAD + PLUMMER_SOURCE * n_sources


def responses(n_subhalos: int) -> int:
    """Base orbit + host Hessian + first-order response vectors (64 subhalos: 3,499)."""
    return plain(0) + HOST_HESSIAN + RESPONSE_PER_SUBHALO * n_subhalos


def mlp(widths=(3, 256, 256, 256, 4)) -> int:
    """Dense MLP forward, 2 FLOP per multiply-accumulate (3-256-256-256-4: 265,728)."""
    return sum(2 * a * b for a, b in zip(widths[:-1], widths[1:]))


def mlp_input_vjp(widths=(3, 256, 256, 256, 4)) -> int:
    """Forward plus the input vector-Jacobian product (the same products transposed), 2 FLOP
    per multiply-accumulate (3-256-256-256-4: 531,456); the per-row work of the surrogate
    adjoint's reverse sweep."""
    return 2 * mlp(widths)


def surrogate() -> int:
    """Surrogate particle-step: MLP in BF16 MFMA plus the FP64 host force and KDK (75)."""
    return plain(0)


def dopri5_step(host_grad: int = HOST_GRAD, n_sources: int = 0) -> int:
    """One Dopri5 step (6 new right-hand sides, FSAL) in a host whose gradient costs
    ``host_grad`` plus n active Plummer sources (3-component host: 774; with a Hernquist
    nucleus: 870; and one active subhalo: 978)."""
    return DOPRI5_STEP + 6 * (host_grad + PLUMMER_SOURCE * n_sources + 6)


def perturb_dopri5_step(n_vectors: int, host_grad: int = HOST_GRAD) -> int:
    """One adaptive step of a base row and n 6-component response vectors (mass and radius
    blocks of n/2 subhalos, or mass blocks of n) under one controller (3-component host:
    1,350 + 576 n; 256 vectors: 148,806)."""
    per_comp = (DOPRI5_STEP - 24) // 6  # 62
    return (DOPRI5_STEP + 6 * (host_grad + HOST_HESSIAN + 6) +
            n_vectors * (6 * per_comp + 6 * RESPONSE_RHS))
