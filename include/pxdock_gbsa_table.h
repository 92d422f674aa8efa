/*
 * pxdock_gbsa_table.h — Born-radius and screening tables and the constants of
 * the generalized-Born / surface-area term of the affinity score.
 *
 * Data that must equal the reference's: kRadiusDatabase / kScreenDatabase
 * (engine/include/bytedock/core/constant.h:234-273, indexed by atomic number,
 * entry 0 the default for elements past the table) and the constructor
 * values of gbsa_total_energy (engine/include/bytedock/core/scorer.h:400-408):
 * the OBC-II rescaling (alpha, beta, gamma) with a 0.09 A radius offset,
 * dielectrics 1 / 78.5, Debye screening kappa, a 1.4 A probe and the
 * surface-tension factor.
 */
#ifndef PXDOCK_GBSA_TABLE_H_
#define PXDOCK_GBSA_TABLE_H_

/* radius by atomic number 0 (default) .. 17 (Cl); an H bonded to N takes
 * PXD_GB_RADIUS_HN (scorer.cpp:742-749) */
#define PXD_GB_NUM_RADII 17
#define PXD_GB_RADII                                                                  \
  { 1.5, 1.2, 1.5, 1.5, 1.5, 1.5, 1.7, 1.55, 1.5, 1.5, 1.5, 1.5, 1.5, 1.5, 2.1, 1.85, \
    1.8, 1.7 }
#define PXD_GB_RADIUS_HN 1.3
/* screening factor by atomic number 0 (default) .. 16 (S) */
#define PXD_GB_NUM_SCREENS 16
#define PXD_GB_SCREENS                                                                  \
  { 0.8, 0.85, 0.8, 0.8, 0.8, 0.8, 0.72, 0.79, 0.85, 0.88, 0.8, 0.8, 0.8, 0.8, 0.8, 0.86, \
    0.96 }

#define PXD_GB_SOLVENT_DIELECTRIC 78.5
#define PXD_GB_SOLUTE_DIELECTRIC 1.0
#define PXD_GB_OFFSET 0.09
#define PXD_GB_ALPHA 1.0
#define PXD_GB_BETA 0.8
#define PXD_GB_GAMMA 4.85
#define PXD_GB_PROBE_RADIUS 1.4
#define PXD_GB_SURFACE_ENERGY 0.0678584012906
#define PXD_GB_KAPPA 0.09293963686178248

#endif /* PXDOCK_GBSA_TABLE_H_ */
