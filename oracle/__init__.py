"""CPU oracle for the flowMC sampling hot path — TEST INFRASTRUCTURE ONLY.

This package restates, in FP64 on the CPU (torch for autograd, numpy for
integer RNG arithmetic), the reference algorithms that the HIP kernels of
``flowmc_amd`` replace. Each function cites the flowMC v0.4.5 file:line it
follows. It is imported only by ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` — as the checker / CPU baseline,
never as a product code path.

Pinning (what independently checks each function; see tests/test_oracle_*.py):
  philox            Random123 known-answer vectors (Philox4x32-10).
  targets           closed forms + central finite differences of autograd
                    gradients; dual-moon known answer -218.14496 from
                    docs/tutorials/dualmoon.ipynb.
  mala / hmc        Gaussian closed forms (linear gradient => exact proposal
                    densities), leapfrog reversibility, acceptance -> 1 as the
                    step -> 0, detailed balance on a 1-D grid.
  nf_mh             target == flow log-prob => every proposal accepted.
  spline / flow     identity at zero logits, fwd(inv(y)) = y, exp(logdet) =
                    finite-difference derivative, change of variables via a
                    full Jacobian, quadrature normalisation in 2-D.
  optim             torch.optim.AdamW + global-norm clip (independent code).
  realnvp           exact inverse, log|det J| from a full autograd Jacobian,
                    2-D quadrature normalisation (= prod sigma, the missing
                    standardisation term), identity at zero output weights,
                    the alternating masks, the reference tests' shapes.
  tempering         exchange: invariance of the product of tempered
                    Gaussians under the swap sweep (moments of every slot),
                    the cold slot receiving a better hot walker, a walker
                    climbing several slots in one sweep, equal temperatures
                    => every swap accepted; adaptation: equal acceptance =>
                    ladder unchanged, a hand-worked 4-rung example, the
                    reference's own [C, T] shaped call; tempered steps: the
                    mala / hmc / grw oracles on log p / T, T = 1 equals the
                    untempered step.
"""
