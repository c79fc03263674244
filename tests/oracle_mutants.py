"""Mutated builds of the CPU oracle, for showing that each pin rejects a
planted mistake (CPU only; nothing here touches a GPU).

Each mutant is a list of (original text, replacement) edits of
oracle/fsmc_oracle.c.  ``build(name, outdir)`` applies them to a copy of the
source -- every original must occur exactly once, so a mutant cannot go stale
silently when the oracle changes -- and compiles it with the oracle's own
flags.  ``oracle.use(path)`` then routes the oracle binding to that build.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
SOURCE = os.path.join(ORACLE_DIR, "fsmc_oracle.c")

# The reference-algorithm LSMC (MonteCarloAmericanOptions/main.cpp:131-279).
_LSMC_REGRESS_BEFORE_DECISION = """    const double* ap = a + (size_t)s * N;
    if (s < n_steps - 1) {
      double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (size_t p = 0; p < N; ++p) {
        const double x = ap[p];
        if (x <= strike) {
          q[0] += 1; q[1] += x; q[2] += x * x; q[3] += x * x * x; q[4] += x * x * x * x;
          q[5] += cf[p]; q[6] += cf[p] * x; q[7] += cf[p] * x * x;
        }
      }
      for (int k = 0; k < 5; ++k) q[k] += 0.01;
      double m[3][3] = {{q[0], q[1], q[2]}, {q[1], q[2], q[3]}, {q[2], q[3], q[4]}}, inv[3][3];
      invert3(inv, m);
      b0 = q[5] * inv[0][0] + q[6] * inv[1][0] + q[7] * inv[2][0];
      b1 = q[5] * inv[0][1] + q[6] * inv[1][1] + q[7] * inv[2][1];
      b2 = q[5] * inv[0][2] + q[6] * inv[1][2] + q[7] * inv[2][2];
    }
"""

MUTANTS = {
    # main.cpp:198 -- the terminal cash flow without its 1/numPaths factor
    "lsmc_no_terminal_scale": [
        ("cf[p] = (strike - x > 0.0 ? strike - x : 0.0) * avgPathFactor;",
         "cf[p] = (strike - x > 0.0 ? strike - x : 0.0);"),
    ],
    # main.cpp:131 -- the sweep starting at numSteps (terminal payoff at
    # maturity) instead of numSteps - 1
    "lsmc_terminal_at_n_steps": [
        ("for (int s = n_steps - 1; s > 0; --s) {", "for (int s = n_steps; s > 0; --s) {"),
        ("if (s < n_steps - 1) {", "if (s < n_steps) {"),
        ("    if (betas) {\n", "    if (betas && s < n_steps) {\n"),  # keep the extra date's fit in bounds
    ],
    # main.cpp:176-190 against :259-261 -- the decision at date s taken with
    # coefficients regressed at date s itself (textbook LSMC) instead of the
    # ones regressed at date s + 1
    "lsmc_no_lag": [
        ("    const double* ap = a + (size_t)s * N;\n", _LSMC_REGRESS_BEFORE_DECISION),
    ],
    # main.cpp:244-246 -- no ridge on the five moments
    "lsmc_no_ridge": [
        ("const double ridge = 0.01;", "const double ridge = 0.0;"),
    ],
    # liborSwaptionGreeks.cpp:70 -- the spot-measure drift term dropped
    "libor_no_drift": [
        ("const double vrat = exp(con1 * v + lam * (sqez - 0.5 * con1));",
         "const double vrat = exp(lam * (sqez - 0.5 * con1));"),
    ],
    # liborSwaptionGreeks.cpp:71 -- the sign of the -1/2 con1 term flipped
    "libor_flipped_ito": [
        ("const double vrat = exp(con1 * v + lam * (sqez - 0.5 * con1));",
         "const double vrat = exp(con1 * v + lam * (sqez + 0.5 * con1));"),
    ],
    # liborSwaptionGreeks.cpp:70 -- the drift with the wrong sign
    "libor_negative_drift": [
        ("const double vrat = exp(con1 * v + lam * (sqez - 0.5 * con1));",
         "const double vrat = exp(-con1 * v + lam * (sqez - 0.5 * con1));"),
    ],
    # MCAsianOptions.cpp:192 -- val2 from the chunk mean (what the sample
    # presumably meant) instead of the running sum
    "heston_val2_chunk_mean": [
        ("val2 += val * val;", "val2 += (tmpVal / L) * (tmpVal / L);"),
    ],
    # MCAsianOptions.cpp:199 -- the quirk flag ignored: per-path standard error
    "heston_val2_ignored": [
        ("if (quirks & ORC_Q_RUNNING_VAL2) {", "if (0) {"),
    ],
    # the discrete geometric average's variance replaced by the continuous
    # average's sigma^2 T / 3
    "geometric_continuous_variance": [
        ("const double var = sigma * sigma * T * (N + 1.0) * (2.0 * N + 1.0) / (6.0 * N * (double)N);",
         "const double var = sigma * sigma * T / 3.0;"),
    ],
    # the discrete average's mean log with the continuous T / 2
    "geometric_continuous_mean": [
        ("const double mu = log(S) + (r - 0.5 * sigma * sigma) * T * (N + 1.0) / (2.0 * N);",
         "const double mu = log(S) + (r - 0.5 * sigma * sigma) * T * 0.5;"),
    ],
}


def mutated_source(name):
    src = open(SOURCE).read()
    for old, new in MUTANTS[name]:
        n = src.count(old)
        if n != 1:
            raise RuntimeError(f"mutant {name}: expected one occurrence of {old!r}, found {n}")
        src = src.replace(old, new)
    return src


def build(name, outdir):
    """Compile mutant ``name`` into outdir; returns the shared library path."""
    os.makedirs(outdir, exist_ok=True)
    src = os.path.join(outdir, f"oracle_{name}.c")
    with open(src, "w") as f:
        f.write(mutated_source(name))
    out = os.path.join(outdir, f"liboracle_{name}.so")
    subprocess.run(["gcc", "-O3", "-std=c99", "-fPIC", "-fopenmp", "-I", ORACLE_DIR, "-shared", "-o", out, src,
                    "-lm"], check=True)
    return out
