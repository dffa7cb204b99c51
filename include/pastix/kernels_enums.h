/*
 * kernels_enums.h -- kernel types of the timing log and of the cost model.
 *
 * The values and their order are those of pastix_ktype_t in the reference
 * (kernels/kernels_enums.h:46-70): model files name the kernels by string and
 * model.csv logs carry the numeric value, so both must agree with it.
 */
#ifndef PASTIX_KERNELS_ENUMS_H
#define PASTIX_KERNELS_ENUMS_H

typedef enum pastix_ktype_e {
    PastixKernelGETRF,          /* LU diagonal block                          */
    PastixKernelHETRF,          /* LDL^H diagonal block                       */
    PastixKernelPOTRF,          /* Cholesky diagonal block                    */
    PastixKernelPXTRF,          /* complex LL^t diagonal block                */
    PastixKernelSYTRF,          /* LDL^T diagonal block                       */
    PastixKernelSCALOCblk,      /* out-of-place scaling of a panel            */
    PastixKernelSCALOBlok,      /* out-of-place scaling of a block            */
    PastixKernelTRSMCblk1d,     /* TRSM of a panel, 1d layout                 */
    PastixKernelTRSMCblk2d,     /* TRSM of a panel, 2d layout                 */
    PastixKernelTRSMCblkLR,     /* TRSM of a panel, low rank                  */
    PastixKernelTRSMBlok2d,     /* TRSM of a block, 2d layout                 */
    PastixKernelTRSMBlokLR,     /* TRSM of a block, low rank                  */
    PastixKernelGEMMCblk1d1d,   /* update panel 1d -> panel 1d                */
    PastixKernelGEMMCblk1d2d,   /* update panel 1d -> panel 2d                */
    PastixKernelGEMMCblk2d2d,   /* update panel 2d -> panel 2d                */
    PastixKernelGEMMCblkFRLR,   /* update full rank -> low rank               */
    PastixKernelGEMMCblkLRLR,   /* update low rank -> low rank                */
    PastixKernelGEMMBlok2d2d,   /* update block 2d -> block 2d                */
    PastixKernelGEMMBlokLRLR,   /* update block low rank -> block low rank    */
    PastixKernelGEADDCblkFRFR,  /* panel addition full rank -> full rank      */
    PastixKernelGEADDCblkFRLR,  /* panel addition full rank -> low rank       */
    PastixKernelGEADDCblkLRLR,  /* panel addition low rank -> low rank        */
    PastixKernelLvl1Nbr
} pastix_ktype_t;

#endif /* PASTIX_KERNELS_ENUMS_H */
