"""Unit system of the reference (galaxy_formation/units.py:17-35): G = 1,
length 3.5 kpc, mass 4e10 Msun, velocity sqrt(G M / L) ~ 222 km/s, time
L / V ~ 15.4 Myr; the cgs constants the sub-grid physics converts with.
Evaluated with the same operations in the same order, so every constant is
bit-identical to the reference's."""

import numpy as np

G_KPC = 4.30091e-6                    # kpc (km/s)^2 / Msun
KPC = 3.5                             # length unit in kpc
MSUN = 4e10                           # mass unit in Msun
KMS = np.sqrt(G_KPC * MSUN / KPC)     # velocity unit in km/s
MYR = KPC / KMS * 977.79              # time unit in Myr

M_H = 1.6726e-24                      # hydrogen mass (g)
K_B = 1.381e-16                       # Boltzmann constant (erg/K)
MSUN_G = 1.989e33                     # solar mass (g)
KPC_CM = 3.086e21                     # kpc (cm)
YR_S = 3.156e7                        # year (s)

RHO_CGS = MSUN * MSUN_G / (KPC * KPC_CM) ** 3     # density unit (g/cm^3)
U_CGS = (KMS * 1e5) ** 2                          # specific energy unit (erg/g)
SEC_PER_MYR = 1e6 * YR_S

X_H = 0.76                            # hydrogen mass fraction
MU_IONISED = 0.6                      # mean molecular weight, ionised primordial gas
