/*
 * CPU restatement of PSDM's 2D acoustic RTM solver -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: it is compiled into liboracle.so and loaded only by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always
 * as the checker.  The product (psdm_amd / libpsdm_rtm.so) never links or calls it.
 *
 * It is included twice from rtm_oracle.c, once with R=float (the reference-faithful
 * FP32 form) and once with R=double (the FP64 twin of the same discrete algorithm).
 * Structure, pass order, clamp rules and indexing follow the reference Fortran on
 * purpose, pass for pass, so that a reader can check it line by line against:
 *   src/solver_ac_forward_ram.f90   (setup :103-319, time loop :344-577)
 *   src/solver_ac_backward_ram.f90  (time loop :329-545, imaging :455-481)
 *   src/solver_ac_forward_disk.f90  (strip re-scatter after the snapshot :506-517)
 *   src/solver_ac_backward_disk.f90 (strip re-scatter after imaging :481-493)
 *   src/solver_space_derivs_clean.f90 (:1-140)
 *   src/solver_PML_matrices.f90 (:1-112)
 *   src/solver_PML_coeffs.f90 (:1-75), src/solver_FD_coeffs.f90 (:1-73)
 * All arrays are column-major (depth index fastest), indices here are 0-based;
 * the Fortran's 1-based index n is n-1 here.
 *
 * Parity pins (tests/): FD tables vs Fornberg weights; PML tables vs the closed
 * form; derivative passes vs brute force; composed solver vs golden vectors
 * written by the reference Fortran compiled from its own sources (oracle/_ref),
 * plus linearity, reciprocity and mirror-symmetry invariants.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

#define AT(a, j, i, ld) (a)[(size_t)(i) * (size_t)(ld) + (size_t)(j)]

/* space_deriv2x, solver_space_derivs_clean.f90:1-37.  k outer, i, j inner.
 * zs = 1: the zero-sum form of the same stencil (not the reference's arithmetic; the
 * opt-in "zero-sum weights" mode of this repository): sum over k = 1..acc of
 * ((p(i+k) - p(i)) + (p(i-k) - p(i))) * coef[k], whose implied centre weight is exactly
 * -2 * (coef[1] + ... + coef[acc]), so a constant field has a zero derivative in any
 * precision.  With FP32 weights the reference form's weights sum to -5.9e-8 instead. */
static void FN(deriv2x)(int acc, const R *coef, int nyt, int nxt, const R *p, R *d, int zs) {
    int i1 = acc, i2 = nxt - 1 - acc;
    memset(d, 0, sizeof(R) * (size_t)nyt * nxt);
    if (zs)
        for (int k = 1; k <= acc; k++)
            for (int i = i1; i <= i2; i++)
                for (int j = 0; j < nyt; j++)
                    AT(d, j, i, nyt) = AT(d, j, i, nyt) + ((AT(p, j, i + k, nyt) - AT(p, j, i, nyt)) +
                                                           (AT(p, j, i - k, nyt) - AT(p, j, i, nyt))) * coef[k];
    else
    for (int k = 0; k <= acc; k++)
        for (int i = i1; i <= i2; i++)
            for (int j = 0; j < nyt; j++)
                AT(d, j, i, nyt) = AT(d, j, i, nyt) + (AT(p, j, i + k, nyt) + AT(p, j, i - k, nyt)) * coef[k];
    for (int i = 0; i < i1; i++)
        for (int j = 0; j < nyt; j++) AT(d, j, i, nyt) = AT(d, j, acc, nyt);
    for (int i = i2 + 1; i < nxt; i++)
        for (int j = 0; j < nyt; j++) AT(d, j, i, nyt) = AT(d, j, nxt - 1 - acc, nyt);
}

/* space_deriv2y, solver_space_derivs_clean.f90:39-69.  i outer, j, k inner.
 * zs = 1: the zero-sum form, as in deriv2x. */
static void FN(deriv2y)(int acc, const R *coef, int nyt, int nxt, const R *p, R *d, int zs) {
    int j1 = acc, j2 = nyt - 1 - acc;
    memset(d, 0, sizeof(R) * (size_t)nyt * nxt);
    for (int i = 0; i < nxt; i++) {
        for (int j = j1; j <= j2; j++)
            if (zs)
                for (int k = 1; k <= acc; k++)
                    AT(d, j, i, nyt) = AT(d, j, i, nyt) + ((AT(p, j + k, i, nyt) - AT(p, j, i, nyt)) +
                                                           (AT(p, j - k, i, nyt) - AT(p, j, i, nyt))) * coef[k];
            else
            for (int k = 0; k <= acc; k++)
                AT(d, j, i, nyt) = AT(d, j, i, nyt) + (AT(p, j + k, i, nyt) + AT(p, j - k, i, nyt)) * coef[k];
        for (int j = 0; j < j1; j++) AT(d, j, i, nyt) = AT(d, acc, i, nyt);
        for (int j = j2 + 1; j < nyt; j++) AT(d, j, i, nyt) = AT(d, nyt - 1 - acc, i, nyt);
    }
}

/* space_derivx, solver_space_derivs_clean.f90:71-107. */
static void FN(derivx)(int acc, const R *coef, int nyt, int nxt, const R *p, R *d) {
    int i1 = acc, i2 = nxt - 1 - acc;
    memset(d, 0, sizeof(R) * (size_t)nyt * nxt);
    for (int k = 0; k <= acc; k++)
        for (int i = i1; i <= i2; i++)
            for (int j = 0; j < nyt; j++)
                AT(d, j, i, nyt) = AT(d, j, i, nyt) + (AT(p, j, i + k, nyt) - AT(p, j, i - k, nyt)) * coef[k];
    for (int i = 0; i < i1; i++)
        for (int j = 0; j < nyt; j++) AT(d, j, i, nyt) = AT(d, j, acc, nyt);
    for (int i = i2 + 1; i < nxt; i++)
        for (int j = 0; j < nyt; j++) AT(d, j, i, nyt) = AT(d, j, nxt - 1 - acc, nyt);
}

/* space_derivy, solver_space_derivs_clean.f90:109-140. */
static void FN(derivy)(int acc, const R *coef, int nyt, int nxt, const R *p, R *d) {
    int j1 = acc, j2 = nyt - 1 - acc;
    memset(d, 0, sizeof(R) * (size_t)nyt * nxt);
    for (int i = 0; i < nxt; i++) {
        for (int j = j1; j <= j2; j++)
            for (int k = 0; k <= acc; k++)
                AT(d, j, i, nyt) = AT(d, j, i, nyt) + (AT(p, j + k, i, nyt) - AT(p, j - k, i, nyt)) * coef[k];
        for (int j = 0; j < j1; j++) AT(d, j, i, nyt) = AT(d, acc, i, nyt);
        for (int j = j2 + 1; j < nyt; j++) AT(d, j, i, nyt) = AT(d, nyt - 1 - acc, i, nyt);
    }
}

/* One CPML block: strip/corner arrays with their own histories
 * (solver_ac_forward_ram.f90:171-220). */
typedef struct {
    int nvy, nvx;    /* block shape */
    int y0, x0;      /* first padded-grid row / column of the block (0-based) */
    R *p1, *p, *p3;  /* the block's own pressure history (xr_pml_p_1, ...) */
    R *phx, *phx1, *chx, *chx1;
    R *phy, *phy1, *chy, *chy1;
    R *prop, *dphi;  /* x_prop_pml / corner_prop_pml and deriv_phi scratch */
    const R *fA, *fB; /* 1D table (strips) or (nvy,nvx) corner table */
} FN(pml_block);

/* pml_matrices_x / _y / _xycorner, solver_PML_matrices.f90:1-112.
 * dir: 0 = x (uses space_derivx on phi), 1 = y (space_derivy).
 * tab: 0 = 1D table indexed by column, 1 = 1D by row, 2 = (nvy,nvx) corner table.
 * rev: 1 reverses the 1D table index (idir=2: factorA(nv-i+1)). */
static void FN(pml_matrices)(int acc, const R *coef1, FN(pml_block) *b, int dir, int tab, int rev,
                             int nyt, const R *deriv, const R *deriv2) {
    R *phi = dir == 0 ? b->phx : b->phy;
    R *phi1 = dir == 0 ? b->phx1 : b->phy1;
    R *chi = dir == 0 ? b->chx : b->chy;
    R *chi1 = dir == 0 ? b->chx1 : b->chy1;
    int nvy = b->nvy, nvx = b->nvx;
    /* deriv_phi of the block's own phi (the value at call time, == phi_1) */
    if (dir == 0) FN(derivx)(acc, coef1, nvy, nvx, phi, b->dphi);
    else FN(derivy)(acc, coef1, nvy, nvx, phi, b->dphi);
    for (int i = 0; i < nvx; i++)
        for (int j = 0; j < nvy; j++) {
            R A, B;
            if (tab == 2) {
                A = AT(b->fA, j, i, nvy);
                B = AT(b->fB, j, i, nvy);
            } else {
                int n = tab == 0 ? nvx : nvy;
                int a = tab == 0 ? i : j;
                if (rev) a = n - 1 - a;
                A = b->fA[a];
                B = b->fB[a];
            }
            R dv = AT(deriv, b->y0 + j, b->x0 + i, nyt);
            R d2 = AT(deriv2, b->y0 + j, b->x0 + i, nyt);
            AT(phi, j, i, nvy) = A * AT(phi1, j, i, nvy) + B * dv;
            AT(chi, j, i, nvy) = A * AT(chi1, j, i, nvy) + B * (d2 + AT(b->dphi, j, i, nvy));
        }
    for (int i = 0; i < nvx * nvy; i++) b->prop[i] = b->dphi[i] + chi[i];
}

/* strip pressure update, e.g. solver_ac_forward_ram.f90:376-377:
 *   p3 = -p1 + 2*p + c2*(d2x + d2y + prop)                                   */
static void FN(block_update)(FN(pml_block) *b, int nyt, const R *c2, const R *d2x, const R *d2y) {
    for (int i = 0; i < b->nvx; i++)
        for (int j = 0; j < b->nvy; j++) {
            int gj = b->y0 + j, gi = b->x0 + i;
            AT(b->p3, j, i, b->nvy) = -AT(b->p1, j, i, b->nvy) + (R)2 * AT(b->p, j, i, b->nvy) +
                                      AT(c2, gj, gi, nyt) * (AT(d2x, gj, gi, nyt) + AT(d2y, gj, gi, nyt) +
                                                            AT(b->prop, j, i, b->nvy));
        }
}

/* corner y-call addition, e.g. :407  lu_pml_p_3 = lu_pml_p_3 + c2*corner_prop_pml */
static void FN(block_add_prop)(FN(pml_block) *b, int nyt, const R *c2) {
    for (int i = 0; i < b->nvx; i++)
        for (int j = 0; j < b->nvy; j++)
            AT(b->p3, j, i, b->nvy) = AT(b->p3, j, i, b->nvy) +
                                      AT(c2, b->y0 + j, b->x0 + i, nyt) * AT(b->prop, j, i, b->nvy);
}

static void FN(block_scatter)(const FN(pml_block) *b, int nyt, R *p3) {
    for (int i = 0; i < b->nvx; i++)
        for (int j = 0; j < b->nvy; j++) AT(p3, b->y0 + j, b->x0 + i, nyt) = AT(b->p3, j, i, b->nvy);
}

/* time actualization of one block (:512-565): p1<-p, p<-p3, phi1<-phi, chi1<-chi */
static void FN(block_rotate)(FN(pml_block) *b) {
    size_t n = (size_t)b->nvx * b->nvy;
    memcpy(b->p1, b->p, n * sizeof(R));
    memcpy(b->p, b->p3, n * sizeof(R));
    memcpy(b->phx1, b->phx, n * sizeof(R));
    memcpy(b->chx1, b->chx, n * sizeof(R));
    memcpy(b->phy1, b->phy, n * sizeof(R));
    memcpy(b->chy1, b->chy, n * sizeof(R));
}

static int FN(block_alloc)(FN(pml_block) *b, int nvy, int nvx, int y0, int x0) {
    size_t n = (size_t)nvy * nvx;
    b->nvy = nvy;
    b->nvx = nvx;
    b->y0 = y0;
    b->x0 = x0;
    R **arrs[] = {&b->p1, &b->p, &b->p3, &b->phx, &b->phx1, &b->chx, &b->chx1,
                  &b->phy, &b->phy1, &b->chy, &b->chy1, &b->prop, &b->dphi};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); a++) {
        *arrs[a] = (R *)calloc(n ? n : 1, sizeof(R));
        if (!*arrs[a]) return -1;
    }
    return 0;
}

static void FN(block_free)(FN(pml_block) *b) {
    R *arrs[] = {b->p1, b->p, b->p3, b->phx, b->phx1, b->chx, b->chx1,
                 b->phy, b->phy1, b->chy, b->chy1, b->prop, b->dphi};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); a++) free(arrs[a]);
}

/*
 * The solver.  iprop = 1 forward (solver_ac_forward_*), 2 backward (solver_ac_backward_*).
 * variant = 0: _ram ordering; 1: _disk ordering (strips scattered again after the
 * free-surface row is zeroed, solver_ac_forward_disk.f90:506-517), which only
 * changes results when fs != 0.
 * table_mode: 0 = coefficient tables computed in FP64 (then rounded to R);
 *             1 = the reference's FP32 expression order (solver_PML_coeffs.f90:16-29).
 * Arrays: S(nt,NumSou) or Sadj(nt,NumRec), ctmp(ny,nx), Fields(ny,nx,nt_snap),
 * Rec_point(nt,NumRec), kernel(ny,nx); geometry indices 1-based as in the reference.
 */
static int FN(solver)(int iprop, int variant, int table_mode, int dPML, int fs, int NumSou, int NumRec,
                      const int *nxSou, const int *nySou, const int *nxRec, const int *nyRec,
                      const float *S, float dmodel_f, float dt_f, int ny, int nx, const float *ctmp,
                      int store_snap, int nt_snap, int nt, R *Fields, R *Rec_point, R *kernel) {
    const int acc = 3;
    const R dmodel = (R)dmodel_f, dt = (R)dt_f;
    /* variant bit 2: "aligned" snapshot pairing for store_snap > 1 -- backward step k
     * images with the snapshot the forward took at step nt + 2 - k.  This is not the
     * reference's bookkeeping (which, with the drivers' nt_snap = nt, reads unfilled
     * slots); it is the pairing the batched engine uses, checked here in FP64. */
    const int aligned = (variant & 2) != 0;
    /* variant bit 4: zero-sum weights in the second-derivative passes (deriv2x above) */
    const int zsum = (variant & 4) != 0;
    variant &= 1;
    int dn = dPML + 2 * acc;
    int nxt = nx + 2 * dn, nyt = dn + ny + dn;
    size_t ng = (size_t)nyt * nxt;
    int rc = 0;
    /* the _ram padding loops read ctmp(1,j) for j up to dn: with nx < dn they read
     * past the end of ctmp, so that case has no defined reference result */
    if (variant == 0 && nx < dn) return -2;
    if (ny < 1 || nx < 1 || nt < 1) return -2;

    int nsrc = iprop == 1 ? NumSou : NumRec;
    const int *sx = iprop == 1 ? nxSou : nxRec;
    const int *sy = iprop == 1 ? nySou : nyRec;

    R *source = (R *)calloc(ng, sizeof(R));
    R *p1 = (R *)calloc(ng, sizeof(R)), *p = (R *)calloc(ng, sizeof(R)), *p3 = (R *)calloc(ng, sizeof(R));
    R *c2 = (R *)calloc(ng, sizeof(R));
    R *d2x = (R *)calloc(ng, sizeof(R)), *d2y = (R *)calloc(ng, sizeof(R));
    R *d1x = (R *)calloc(ng, sizeof(R)), *d1y = (R *)calloc(ng, sizeof(R));
    R *fA = (R *)calloc(dn, sizeof(R)), *fB = (R *)calloc(dn, sizeof(R));
    R *cA[4], *cB[4]; /* ld, rd, lu, ru */
    for (int q = 0; q < 4; q++) {
        cA[q] = (R *)calloc((size_t)dn * dn, sizeof(R));
        cB[q] = (R *)calloc((size_t)dn * dn, sizeof(R));
    }
    R coef1[4], coef2[4];
    FN(pml_block) xl, xr, yu, yd, lu, ru, ld, rd;
    FN(pml_block) *blocks[8] = {&xl, &xr, &yu, &yd, &lu, &ru, &ld, &rd};
    memset(blocks[0], 0, sizeof(xl));
    for (int q = 0; q < 8; q++) memset(blocks[q], 0, sizeof(xl));
    rc |= FN(block_alloc)(&xr, ny, dn, dn, dn + nx);
    rc |= FN(block_alloc)(&xl, ny, dn, dn, 0);
    rc |= FN(block_alloc)(&yd, dn, nx, dn + ny, dn);
    rc |= FN(block_alloc)(&yu, dn, nx, 0, dn);
    rc |= FN(block_alloc)(&lu, dn, dn, 0, 0);
    rc |= FN(block_alloc)(&ru, dn, dn, 0, dn + nx);
    rc |= FN(block_alloc)(&ld, dn, dn, dn + ny, 0);
    rc |= FN(block_alloc)(&rd, dn, dn, dn + ny, dn + nx);
    if (rc || !source || !p1 || !p || !p3 || !c2 || !d2x || !d2y || !d1x || !d1y) {
        rc = -1;
        goto done;
    }

    /* velocity padding then c2 (:319).  c starts at zero (:152).
     * _ram (solver_ac_forward_ram.f90:282-309): the top and bottom pads are filled by
     *   loops over j=1,dnx only, so just the first dn interior columns get ctmp(1,j) /
     *   ctmp(ny,j); the rest of those pads keep c=0 (and so c2=0).
     * _disk (solver_ac_forward_disk.f90:294-313): whole rows, c(i,dn+1:dn+nx)=ctmp(1,:). */
    {
        R *c = d2x; /* scratch */
        int ntop = variant == 0 ? dn : nx; /* interior columns that get a top/bottom pad value */
        memset(c, 0, ng * sizeof(R));
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) AT(c, dn + j, dn + i, nyt) = (R)AT(ctmp, j, i, ny);
        for (int i = 0; i < dn; i++)
            for (int j = 0; j < dn; j++) {
                AT(c, j, i, nyt) = (R)AT(ctmp, 0, 0, ny);
                AT(c, j, dn + nx + i, nyt) = (R)AT(ctmp, 0, nx - 1, ny);
                AT(c, dn + ny + j, i, nyt) = (R)AT(ctmp, ny - 1, 0, ny);
                AT(c, dn + ny + j, dn + nx + i, nyt) = (R)AT(ctmp, ny - 1, nx - 1, ny);
            }
        for (int i = 0; i < ntop; i++)
            for (int j = 0; j < dn; j++) {
                AT(c, j, dn + i, nyt) = (R)AT(ctmp, 0, i, ny);
                AT(c, dn + ny + j, dn + i, nyt) = (R)AT(ctmp, ny - 1, i, ny);
            }
        for (int i = 0; i < dn; i++)
            for (int j = 0; j < ny; j++) {
                AT(c, dn + j, i, nyt) = (R)AT(ctmp, j, 0, ny);
                AT(c, dn + j, dn + nx + i, nyt) = (R)AT(ctmp, j, nx - 1, ny);
            }
        for (size_t n = 0; n < ng; n++) c2[n] = ((c[n] * c[n]) * (dt * dt)) / (dmodel * dmodel);
    }

    /* PML_coeffs (solver_PML_coeffs.f90), finite_difference_coef (solver_FD_coeffs.f90) */
    FN(pml_tables)(table_mode, acc, dPML, dt_f, dmodel_f, dn, fA, fB, cA[0], cA[1], cA[2], cA[3], cB[0], cB[1],
                   cB[2], cB[3]);
    FN(fd_coef)(1, coef1);
    FN(fd_coef)(2, coef2);
    xr.fA = fA; xr.fB = fB;
    xl.fA = fA; xl.fB = fB;
    yd.fA = fA; yd.fB = fB;
    yu.fA = fA; yu.fB = fB;
    ld.fA = cA[0]; ld.fB = cB[0];
    rd.fA = cA[1]; rd.fB = cB[1];
    lu.fA = cA[2]; lu.fB = cB[2];
    ru.fA = cA[3]; ru.fB = cB[3];

    if (iprop == 1 && Rec_point) memset(Rec_point, 0, sizeof(R) * (size_t)nt * NumRec);

    int l = 0;
    int snap = iprop == 1 ? 0 : nt_snap;
    /* snapshot index s (1-based in the reference) lives at Fields(:,:,s) */
    for (int k = 2; k <= nt; k++) {
        l++;
        /* injection (:348-358): last writer wins */
        for (int s = 0; s < nsrc; s++) {
            int jj = dn + sy[s] - 1, ii = dn + sx[s] - 1;
            if (jj >= 0 && jj < nyt && ii >= 0 && ii < nxt)
                AT(source, jj, ii, nyt) = (R)S[(size_t)s * nt + (k - 2)];
        }
        FN(deriv2x)(acc, coef2, nyt, nxt, p, d2x, zsum);
        FN(deriv2y)(acc, coef2, nyt, nxt, p, d2y, zsum);
        FN(derivx)(acc, coef1, nyt, nxt, p, d1x);
        FN(derivy)(acc, coef1, nyt, nxt, p, d1y);

        /* x strips (:374-382) */
        FN(pml_matrices)(acc, coef1, &xr, 0, 0, 0, nyt, d1x, d2x);
        FN(block_update)(&xr, nyt, c2, d2x, d2y);
        FN(pml_matrices)(acc, coef1, &xl, 0, 0, 1, nyt, d1x, d2x);
        FN(block_update)(&xl, nyt, c2, d2x, d2y);
        /* y strips (:386-396) */
        FN(pml_matrices)(acc, coef1, &yd, 1, 1, 0, nyt, d1y, d2y);
        FN(block_update)(&yd, nyt, c2, d2x, d2y);
        if (fs == 0) {
            FN(pml_matrices)(acc, coef1, &yu, 1, 1, 1, nyt, d1y, d2y);
            FN(block_update)(&yu, nyt, c2, d2x, d2y);
            /* corners up (:400-416) */
            FN(pml_matrices)(acc, coef1, &lu, 0, 2, 0, nyt, d1x, d2x);
            FN(block_update)(&lu, nyt, c2, d2x, d2y);
            FN(pml_matrices)(acc, coef1, &lu, 1, 2, 0, nyt, d1y, d2y);
            FN(block_add_prop)(&lu, nyt, c2);
            FN(pml_matrices)(acc, coef1, &ru, 0, 2, 0, nyt, d1x, d2x);
            FN(block_update)(&ru, nyt, c2, d2x, d2y);
            FN(pml_matrices)(acc, coef1, &ru, 1, 2, 0, nyt, d1y, d2y);
            FN(block_add_prop)(&ru, nyt, c2);
        }
        /* corners down (:422-438) */
        FN(pml_matrices)(acc, coef1, &ld, 0, 2, 0, nyt, d1x, d2x);
        FN(block_update)(&ld, nyt, c2, d2x, d2y);
        FN(pml_matrices)(acc, coef1, &ld, 1, 2, 0, nyt, d1y, d2y);
        FN(block_add_prop)(&ld, nyt, c2);
        FN(pml_matrices)(acc, coef1, &rd, 0, 2, 0, nyt, d1x, d2x);
        FN(block_update)(&rd, nyt, c2, d2x, d2y);
        FN(pml_matrices)(acc, coef1, &rd, 1, 2, 0, nyt, d1y, d2y);
        FN(block_add_prop)(&rd, nyt, c2);

        /* interior (:443-444): p3 = -p1 + 2p + (d2x + d2y + source)*c2 */
        for (int i = dn; i < dn + nx; i++)
            for (int j = dn; j < dn + ny; j++)
                AT(p3, j, i, nyt) = -AT(p1, j, i, nyt) + (R)2 * AT(p, j, i, nyt) +
                                    (AT(d2x, j, i, nyt) + AT(d2y, j, i, nyt) + AT(source, j, i, nyt)) *
                                        AT(c2, j, i, nyt);
        /* scatter (:449-459) */
        for (int q = 0; q < 8; q++) FN(block_scatter)(blocks[q], nyt, p3);
        /* free surface (:463-473) */
        if (fs != 0)
            for (int i = 0; i < nxt; i++) AT(p3, dn, i, nyt) = 0;

        if (aligned && iprop == 2) {
            if ((nt + 1 - k) % store_snap == 0) {
                int s = (nt + 1 - k) / store_snap;
                if (kernel && Fields && s >= 1 && s <= nt_snap)
                    for (int i = 0; i < nx; i++)
                        for (int j = 0; j < ny; j++)
                            AT(kernel, j, i, ny) =
                                AT(kernel, j, i, ny) + (R)0.5 * (AT(p3, dn + j, dn + i, nyt) - AT(p1, dn + j, dn + i, nyt)) *
                                                           Fields[((size_t)(s - 1) * nx + i) * ny + j] * (R)store_snap;
            }
            l = 0;
        } else if (l == store_snap) {
            if (iprop == 1) {
                snap++;
                if (Fields && snap >= 1 && snap <= nt_snap)
                    for (int i = 0; i < nx; i++)
                        for (int j = 0; j < ny; j++)
                            Fields[((size_t)(snap - 1) * nx + i) * ny + j] =
                                ((AT(p3, dn + j, dn + i, nyt) - AT(p1, dn + j, dn + i, nyt)) * (R)0.5) / dt;
            } else {
                snap--;
                if (kernel && Fields && snap >= 1 && snap <= nt_snap)
                    for (int i = 0; i < nx; i++)
                        for (int j = 0; j < ny; j++)
                            AT(kernel, j, i, ny) =
                                AT(kernel, j, i, ny) + (R)0.5 * (AT(p3, dn + j, dn + i, nyt) - AT(p1, dn + j, dn + i, nyt)) *
                                                           Fields[((size_t)(snap - 1) * nx + i) * ny + j] * (R)store_snap;
            }
            l = 0;
        }
        /* _disk ordering: strips are scattered again after the snapshot/imaging */
        if (variant == 1)
            for (int q = 0; q < 8; q++) FN(block_scatter)(blocks[q], nyt, p3);

        /* time actualization (:509-565) */
        memcpy(p1, p, ng * sizeof(R));
        memcpy(p, p3, ng * sizeof(R));
        for (int q = 0; q < 8; q++) FN(block_rotate)(blocks[q]);

        /* traces (:567-571) */
        if (iprop == 1 && Rec_point)
            for (int r = 0; r < NumRec; r++) {
                int jj = dn + nyRec[r] - 1, ii = dn + nxRec[r] - 1;
                R v = (jj >= 0 && jj < nyt && ii >= 0 && ii < nxt) ? AT(p3, jj, ii, nyt) : (R)0;
                Rec_point[(size_t)r * nt + (k - 2)] = v;
            }
    }
    if (iprop == 1 && Rec_point && nt >= 2)
        for (int r = 0; r < NumRec; r++) Rec_point[(size_t)r * nt + nt - 1] = Rec_point[(size_t)r * nt + nt - 2];

done:
    for (int q = 0; q < 8; q++) FN(block_free)(blocks[q]);
    for (int q = 0; q < 4; q++) {
        free(cA[q]);
        free(cB[q]);
    }
    free(fA); free(fB);
    free(source); free(p1); free(p); free(p3); free(c2);
    free(d2x); free(d2y); free(d1x); free(d1y);
    return rc;
}

#undef FN
#undef AT
