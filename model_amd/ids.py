"""Field, role and column identifiers of the C-ABI (include/dyncore.h)."""

# prognostic, diagnostic and basic-state fields (model.h:63-91, microphysics.h:18-51)
U, V, W, TH, QV, QC, QR, PI, UBAR, VBAR, WBAR, THBAR, QBAR, PBAR, VT, ACCRAIN = range(16)
FIELD_NAMES = ["u", "v", "w", "th", "qv", "qc", "qr", "pi", "ubar", "vbar", "wbar", "thbar",
               "qbar", "pbar", "vt", "accrain"]
PROGNOSTIC = (U, V, W, TH, QV, QC, QR)
BASIC = (UBAR, VBAR, WBAR, THBAR, QBAR, PBAR)

# the three rotating buffers of a prognostic: the reference's ums/us/ups (solver.cpp:229-263)
ROLE_M, ROLE_CUR, ROLE_P = 0, 1, 2

# 1-D columns (model.h:231-245); F is the Coriolis parameter f[j] of length NY
(COL_ZU, COL_ZW, COL_ZSU, COL_ZSW, COL_MU, COL_MW, COL_RHOU, COL_RHOW, COL_ONE_D_RHOU,
 COL_ONE_D_RHOW, COL_TB, COL_TBV, COL_TBW, COL_QB, COL_PIB, COL_F) = range(16)
COLUMN_NAMES = ["zu", "zw", "zsu", "zsw", "mu", "mw", "rhou", "rhow", "one_d_rhou", "one_d_rhow",
                "tb", "tbv", "tbw", "qb", "pib", "f"]
COLUMN_ID = {n: c for c, n in enumerate(COLUMN_NAMES)}
FIELD_ID = {n: f for f, n in enumerate(FIELD_NAMES)}
