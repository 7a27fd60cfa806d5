/*
 * orc_impl.h -- real-type-generic part of the oracle (TEST INFRASTRUCTURE).
 *
 * Included twice by fgs_oracle.c: with R=float (suffix _f32, FP32 emulation,
 * same operation order as the GPU kernels, no contraction) and with R=double
 * (suffix _f64).  Matrices follow glm's column-major convention M[col][row];
 * every product is written in the order glm evaluates it
 * (antialias/third_party/glm/glm/detail/type_mat3x3.inl:486-518:
 * Result[j][i] = m1[0][i]*m2[j][0] + m1[1][i]*m2[j][1] + m1[2][i]*m2[j][2]).
 *
 * Pinning (checks in tests/test_oracle_*.py that do not restate these
 * formulas):
 *   preprocess   : cov3D vs scipy Rotation (quaternion convention), Sigma2D vs
 *                  the numerically differentiated pinhole projection J*Sigma*J^T,
 *                  means2D vs fx*X/Z + cx - 0.5, conic*Sigma2D = I, OGFS radius
 *                  bounds and closed forms, SH basis orthonormality by
 *                  quadrature, SH degree-0 closed form.
 *   blend fwd    : single-Gaussian closed form, empty tile = bg, early-stop
 *                  invariant T >= 1e-4, brute-force per-pixel depth-sorted
 *                  compositing with numpy cumprod.
 *   blend bwd +
 *   preprocess bwd: FP64 central finite differences of the f64 forward
 *                  (grad_mode exact), and reference == exact where the two
 *                  coincide (dilate variant, no alpha clamp, mod = 1); the
 *                  reference mode's cov2D epsilon by the ratio
 *                  reference / exact = denom^2 / (denom^2 + 1e-7); the
 *                  alpha clamp by the closed form of one Gaussian with
 *                  o*G > 0.99 on a pixel centre (reference mode passes
 *                  dL/do = G (c - bg) dL/dpix, exact mode 0).
 *   nthreads > 1 : blend fwd/bwd split tiles over threads (per-thread
 *                  accumulators reduced in thread order), preprocess fwd/bwd
 *                  split Gaussians; nthreads = 1 keeps the sequential order.
 */

#define CF(x) ((R)(x##f))

/* transformPoint4x4 / transformPoint4x3 (auxiliary.h:58-77). */
static inline void SFX(xf4)(const float* m, R px, R py, R pz, R* o)
{
    o[0] = (R)m[0] * px + (R)m[4] * py + (R)m[8] * pz + (R)m[12];
    o[1] = (R)m[1] * px + (R)m[5] * py + (R)m[9] * pz + (R)m[13];
    o[2] = (R)m[2] * px + (R)m[6] * py + (R)m[10] * pz + (R)m[14];
    o[3] = (R)m[3] * px + (R)m[7] * py + (R)m[11] * pz + (R)m[15];
}
static inline void SFX(xf3)(const float* m, R px, R py, R pz, R* o)
{
    o[0] = (R)m[0] * px + (R)m[4] * py + (R)m[8] * pz + (R)m[12];
    o[1] = (R)m[1] * px + (R)m[5] * py + (R)m[9] * pz + (R)m[13];
    o[2] = (R)m[2] * px + (R)m[6] * py + (R)m[10] * pz + (R)m[14];
}

/* Rotation matrix of the (unnormalised) quaternion q = (r, x, y, z) in glm
 * column-major form, forward.cu:167-171. */
static inline void SFX(quat_rot)(R r, R x, R y, R z, R Rm[3][3])
{
    Rm[0][0] = CF(1.0) - CF(2.0) * (y * y + z * z);
    Rm[0][1] = CF(2.0) * (x * y - r * z);
    Rm[0][2] = CF(2.0) * (x * z + r * y);
    Rm[1][0] = CF(2.0) * (x * y + r * z);
    Rm[1][1] = CF(1.0) - CF(2.0) * (x * x + z * z);
    Rm[1][2] = CF(2.0) * (y * z - r * x);
    Rm[2][0] = CF(2.0) * (x * z - r * y);
    Rm[2][1] = CF(2.0) * (y * z + r * x);
    Rm[2][2] = CF(1.0) - CF(2.0) * (x * x + y * y);
}

/* computeCov3D (forward.cu:151-185): Sigma = (S R)^T (S R), upper 6. */
static inline void SFX(cov3d)(R sx, R sy, R sz, R r, R x, R y, R z, R* cov)
{
    R Rm[3][3], Mm[3][3];
    SFX(quat_rot)(r, x, y, z, Rm);
    const R s[3] = { sx, sy, sz };
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            Mm[j][i] = s[i] * Rm[j][i];
#define SIG(j, i) (Mm[i][0] * Mm[j][0] + Mm[i][1] * Mm[j][1] + Mm[i][2] * Mm[j][2])
    cov[0] = SIG(0, 0);
    cov[1] = SIG(0, 1);
    cov[2] = SIG(0, 2);
    cov[3] = SIG(1, 1);
    cov[4] = SIG(1, 2);
    cov[5] = SIG(2, 2);
#undef SIG
}

/* EWA 2D covariance (forward.cu:102-146), before any dilation.  Also returns
 * the clamped camera-space mean, J entries and T = W*J for reuse. */
typedef struct {
    R t[3];
    R txtz, tytz, limx, limy;
    R J00, J02, J11, J12;
    R W[3][3];
    R T[2][3];
    R A[3][2];
    R a, b, c;
} SFX(cov2d_t);

static inline void SFX(cov2d)(const float* V, R px, R py, R pz, R fx, R fy, R tanx, R tany,
                              const R* cov3, SFX(cov2d_t)* o)
{
    SFX(xf3)(V, px, py, pz, o->t);
    o->limx = CF(1.3) * tanx;
    o->limy = CF(1.3) * tany;
    o->txtz = o->t[0] / o->t[2];
    o->tytz = o->t[1] / o->t[2];
    o->t[0] = RMIN(o->limx, RMAX(-o->limx, o->txtz)) * o->t[2];
    o->t[1] = RMIN(o->limy, RMAX(-o->limy, o->tytz)) * o->t[2];
    const R tz = o->t[2];
    o->J00 = fx / tz;
    o->J02 = -(fx * o->t[0]) / (tz * tz);
    o->J11 = fy / tz;
    o->J12 = -(fy * o->t[1]) / (tz * tz);
    /* W = mat3(V0,V4,V8, V1,V5,V9, V2,V6,V10): W[c][r] */
    o->W[0][0] = V[0]; o->W[0][1] = V[4]; o->W[0][2] = V[8];
    o->W[1][0] = V[1]; o->W[1][1] = V[5]; o->W[1][2] = V[9];
    o->W[2][0] = V[2]; o->W[2][1] = V[6]; o->W[2][2] = V[10];
    /* T = W * J with J[0] = (J00,0,J02), J[1] = (0,J11,J12), J[2] = 0 */
    for (int i = 0; i < 3; ++i) {
        o->T[0][i] = o->W[0][i] * o->J00 + o->W[2][i] * o->J02;
        o->T[1][i] = o->W[1][i] * o->J11 + o->W[2][i] * o->J12;
    }
    /* Vrk columns */
    const R Vk[3][3] = { { cov3[0], cov3[1], cov3[2] },
                         { cov3[1], cov3[3], cov3[4] },
                         { cov3[2], cov3[4], cov3[5] } };
    /* A = T^T * Vrk : A[j][i] = T[i][0]*Vk[j][0] + T[i][1]*Vk[j][1] + T[i][2]*Vk[j][2] */
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 2; ++i)
            o->A[j][i] = o->T[i][0] * Vk[j][0] + o->T[i][1] * Vk[j][1] + o->T[i][2] * Vk[j][2];
    /* cov = A * T (T[2] = 0) */
    o->a = o->A[0][0] * o->T[0][0] + o->A[1][0] * o->T[0][1] + o->A[2][0] * o->T[0][2];
    o->b = o->A[0][1] * o->T[0][0] + o->A[1][1] * o->T[0][1] + o->A[2][1] * o->T[0][2];
    o->c = o->A[0][1] * o->T[1][0] + o->A[1][1] * o->T[1][1] + o->A[2][1] * o->T[1][2];
}

/* getRect (auxiliary.h:46-56): C truncating division by the 16-pixel tile,
 * clamped to [0, grid]; evaluated in R (float in the f32 mode). */
static inline void SFX(get_rect)(R px, R py, int r, int gx, int gy, int* out)
{
    const R rr = (R)r;
    int x0 = orc_sat_int((double)((px - rr) / CF(16.0)));
    int y0 = orc_sat_int((double)((py - rr) / CF(16.0)));
    int x1 = orc_sat_int((double)((px + rr + CF(16.0) - CF(1.0)) / CF(16.0)));
    int y1 = orc_sat_int((double)((py + rr + CF(16.0) - CF(1.0)) / CF(16.0)));
    out[0] = orc_clampi(x0, 0, gx);
    out[1] = orc_clampi(y0, 0, gy);
    out[2] = orc_clampi(x1, 0, gx);
    out[3] = orc_clampi(y1, 0, gy);
}

/* SH constants (auxiliary.h:22-39). */
static const R SFX(SHC0) = CF(0.28209479177387814);
static const R SFX(SHC1) = CF(0.4886025119029199);
static const R SFX(SHC2)[5] = { CF(1.0925484305920792), CF(-1.0925484305920792),
                                CF(0.31539156525252005), CF(-1.0925484305920792),
                                CF(0.5462742152960396) };
static const R SFX(SHC3)[7] = { CF(-0.5900435899266435), CF(2.890611442640554),
                                CF(-0.4570457994644658), CF(0.3731763325901154),
                                CF(-0.4570457994644658), CF(1.445305721320277),
                                CF(-0.5900435899266435) };

/* computeColorFromSH (forward.cu:48-99) for one channel c; sh points at the
 * Gaussian's M*3 coefficients. Returns the unclamped value (before +0.5 is
 * NOT applied: the caller adds 0.5). */
static inline R SFX(sh_eval)(int deg, const R* sh, int c, R x, R y, R z)
{
#define SH(k) sh[(k) * 3 + c]
    R res = SFX(SHC0) * SH(0);
    if (deg > 0) {
        res = res - SFX(SHC1) * y * SH(1) + SFX(SHC1) * z * SH(2) - SFX(SHC1) * x * SH(3);
        if (deg > 1) {
            R xx = x * x, yy = y * y, zz = z * z;
            R xy = x * y, yz = y * z, xz = x * z;
            res = res + SFX(SHC2)[0] * xy * SH(4) + SFX(SHC2)[1] * yz * SH(5) +
                  SFX(SHC2)[2] * (CF(2.0) * zz - xx - yy) * SH(6) + SFX(SHC2)[3] * xz * SH(7) +
                  SFX(SHC2)[4] * (xx - yy) * SH(8);
            if (deg > 2) {
                res = res + SFX(SHC3)[0] * y * (CF(3.0) * xx - yy) * SH(9) +
                      SFX(SHC3)[1] * xy * z * SH(10) +
                      SFX(SHC3)[2] * y * (CF(4.0) * zz - xx - yy) * SH(11) +
                      SFX(SHC3)[3] * z * (CF(2.0) * zz - CF(3.0) * xx - CF(3.0) * yy) * SH(12) +
                      SFX(SHC3)[4] * x * (CF(4.0) * zz - xx - yy) * SH(13) +
                      SFX(SHC3)[5] * z * (xx - yy) * SH(14) +
                      SFX(SHC3)[6] * x * (xx - CF(3.0) * yy) * SH(15);
            }
        }
    }
#undef SH
    return res;
}

static inline void SFX(view_dir)(const float* campos, R px, R py, R pz, R* d, R* dir_orig)
{
    dir_orig[0] = px - (R)campos[0];
    dir_orig[1] = py - (R)campos[1];
    dir_orig[2] = pz - (R)campos[2];
    R len = RSQRT(dir_orig[0] * dir_orig[0] + dir_orig[1] * dir_orig[1] + dir_orig[2] * dir_orig[2]);
    d[0] = dir_orig[0] / len;
    d[1] = dir_orig[1] / len;
    d[2] = dir_orig[2] / len;
}

/* ---- preprocess forward (forward.cu:266-401) ---------------------------- */
void SFX(orc_preprocess)(const orc_settings* s, int32_t P, int32_t M,
                         const R* means3D, const R* scales, const R* rotations,
                         const R* opacities, const R* shs, const R* colors_precomp,
                         const R* cov3D_precomp,
                         int32_t* radii, R* means2D, R* depths, R* cov3D, R* rgb,
                         R* conic_opacity, uint32_t* tiles_touched, uint8_t* clamped,
                         int32_t* rect, int32_t nthreads)
{
    const int W = s->image_width, H = s->image_height;
    const float focal_y_f = (float)H / (2.0f * s->tanfovy);   /* rasterizer_impl.cu:283-284 */
    const float focal_x_f = (float)W / (2.0f * s->tanfovx);
    const R fx = (R)focal_x_f, fy = (R)focal_y_f;
    const R tanx = (R)s->tanfovx, tany = (R)s->tanfovy;
    const int gx = (W + 15) / 16, gy = (H + 15) / 16;
    const float log_lambda_f = orc_ogfs_log_lambda(s->ogfs_epsilon_scale, s->ogfs_epsilon_max);
    const R log_lambda = (R)log_lambda_f;
    const R mod = (R)s->scale_modifier;
    const int deg = s->sh_degree;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = 0; i < P; ++i) {
        radii[i] = 0;
        tiles_touched[i] = 0;
        means2D[2 * i] = 0; means2D[2 * i + 1] = 0;
        depths[i] = 0;
        for (int k = 0; k < 4; ++k) conic_opacity[4 * i + k] = 0;
        for (int k = 0; k < 3; ++k) { rgb[3 * i + k] = 0; clamped[3 * i + k] = 0; }
        for (int k = 0; k < 6; ++k) cov3D[6 * i + k] = 0;
        for (int k = 0; k < 4; ++k) rect[4 * i + k] = 0;

        const R px = means3D[3 * i], py = means3D[3 * i + 1], pz = means3D[3 * i + 2];
        R hom[4], view[3];
        SFX(xf4)(s->projmatrix, px, py, pz, hom);
        const R p_w = CF(1.0) / (hom[3] + CF(0.0000001));
        const R projx = hom[0] * p_w, projy = hom[1] * p_w;
        SFX(xf3)(s->viewmatrix, px, py, pz, view);
        if (view[2] <= CF(0.2))                       /* in_frustum, auxiliary.h:154 */
            continue;

        R c3[6];
        if (cov3D_precomp) {
            for (int k = 0; k < 6; ++k) c3[k] = cov3D_precomp[6 * i + k];
        } else {
            SFX(cov3d)(mod * scales[3 * i], mod * scales[3 * i + 1], mod * scales[3 * i + 2],
                       rotations[4 * i], rotations[4 * i + 1], rotations[4 * i + 2],
                       rotations[4 * i + 3], c3);
        }
        for (int k = 0; k < 6; ++k) cov3D[6 * i + k] = c3[k];

        SFX(cov2d_t) cv;
        SFX(cov2d)(s->viewmatrix, px, py, pz, fx, fy, tanx, tany, c3, &cv);
        R a = cv.a, b = cv.b, c = cv.c;
        if (s->variant == 1) {                    /* dilate: += DILATE_PIXEL (double) */
            a = (R)((double)a + 0.3);
            c = (R)((double)c + 0.3);
        } else if (s->use_filter) {               /* wodilate: max(., 0.3), forward.cu:141-144 */
            a = RMAX(a, CF(0.3));
            c = RMAX(c, CF(0.3));
        }
        const R det = a * c - b * b;
        if (det == CF(0.0))
            continue;
        const R det_inv = CF(1.0) / det;
        const R con0 = c * det_inv, con1 = -b * det_inv, con2 = a * det_inv;

        int radius_px;
        R opacity;
        if (s->variant == 1) {
            /* upstream 3DGS radius: ceil(3 sqrt(max(l1, l2))), disc clamp 0.1 */
            opacity = opacities[i];
            const R mid = CF(0.5) * (a + c);
            const R l1 = mid + RSQRT(RMAX(CF(0.1), mid * mid - det));
            const R l2 = mid - RSQRT(RMAX(CF(0.1), mid * mid - det));
            radius_px = orc_sat_int((double)RCEIL(CF(3.0) * RSQRT(RMAX(l1, l2))));
            if (radius_px <= 0)
                continue;
        } else {
            /* OGFS energy radius, forward.cu:345-369 */
            opacity = RMAX(opacities[i], CF(1e-6));
            const R log_opacity = RLOG(opacity);
            const R radius_energy = log_opacity - log_lambda;
            const R energy_clamped = RMAX(radius_energy, (R)s->ogfs_energy_floor);
            const R mid = CF(0.5) * (a + c);
            const R disc = RMAX(CF(0.0), mid * mid - det);
            const R sqrt_disc = RSQRT(disc);
            const R l1 = mid + sqrt_disc, l2 = mid - sqrt_disc;
            const R max_lambda = RMAX(RMAX(l1, l2), CF(1e-6));
            const R radius_base = CF(3.0) * RSQRT(max_lambda);
            R radius_major = RSQRT(CF(2.0) * energy_clamped * max_lambda);
            radius_major = RMIN(radius_major, radius_base);
            radius_major = RMAX(radius_major, (R)s->ogfs_radius_ratio_min * radius_base);
            const R radius_f = radius_major + CF(1.0);
            int rp = orc_sat_int((double)RCEIL(radius_f));
            radius_px = rp > 1 ? rp : 1;
        }
        /* ndc2Pix in double (auxiliary.h:41-44) */
        const R mx = (R)((((double)projx + 1.0) * (double)W - 1.0) * 0.5);
        const R my = (R)((((double)projy + 1.0) * (double)H - 1.0) * 0.5);
        int rc[4];
        SFX(get_rect)(mx, my, radius_px, gx, gy, rc);
        const int tiles = (rc[2] - rc[0]) * (rc[3] - rc[1]);
        if (tiles == 0)
            continue;

        if (!colors_precomp) {
            R d[3], dor[3];
            SFX(view_dir)(s->campos, px, py, pz, d, dor);
            for (int ch = 0; ch < 3; ++ch) {
                R v = SFX(sh_eval)(deg, shs + (size_t)i * M * 3, ch, d[0], d[1], d[2]);
                v += CF(0.5);
                clamped[3 * i + ch] = v < CF(0.0);
                rgb[3 * i + ch] = (v < CF(0.0)) ? CF(0.0) : v;
            }
        } else {
            for (int ch = 0; ch < 3; ++ch) rgb[3 * i + ch] = colors_precomp[3 * i + ch];
        }
        depths[i] = view[2];
        radii[i] = radius_px;
        means2D[2 * i] = mx;
        means2D[2 * i + 1] = my;
        conic_opacity[4 * i + 0] = con0;
        conic_opacity[4 * i + 1] = con1;
        conic_opacity[4 * i + 2] = con2;
        conic_opacity[4 * i + 3] = opacity;
        tiles_touched[i] = (uint32_t)tiles;
        for (int k = 0; k < 4; ++k) rect[4 * i + k] = rc[k];
    }
}

/* ---- blend forward (forward.cu:406-567) --------------------------------- */
void SFX(orc_blend_fwd)(const orc_settings* s, int32_t P, const uint32_t* ranges,
                        const uint32_t* point_list, const R* means2D, const R* colors,
                        const R* conic_opacity, R* out_color, R* final_T,
                        uint32_t* n_contrib, int32_t* point_id, R* weight_pixel,
                        R* point_weight, R* kv_alpha_sum, uint32_t* kv_alpha_count,
                        uint64_t* decision_hash, int32_t nthreads)
{
    const int W = s->image_width, H = s->image_height;
    const int gx = (W + 15) / 16, gy = (H + 15) / 16;
    const int T_ = gx * gy;
    int nt = nthreads > 0 ? nthreads : 1;
    /* per-thread private point_weight maxima, reduced at the end */
    R* pw_priv = (R*)calloc((size_t)nt * (size_t)(P > 0 ? P : 1), sizeof(R));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
#endif
    for (int tile = 0; tile < T_; ++tile) {
#ifdef _OPENMP
        R* pw = pw_priv + (size_t)omp_get_thread_num() * (size_t)P;
#else
        R* pw = pw_priv;
#endif
        const int tx = tile % gx, ty = tile / gx;
        const uint32_t r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
        for (int ly = 0; ly < 16; ++ly) {
            for (int lx = 0; lx < 16; ++lx) {
                const int pxi = tx * 16 + lx, pyi = ty * 16 + ly;
                if (pxi >= W || pyi >= H)
                    continue;
                const int pix_id = W * pyi + pxi;
                const R pfx = (R)pxi, pfy = (R)pyi;
                R T = CF(1.0);
                R C[3] = { 0, 0, 0 };
                uint32_t contributor = 0, last_contributor = 0;
                R weight_max = CF(0.0);
                int max_id = -1;
                uint64_t h = 1469598103934665603ull;
                for (uint32_t k = r0; k < r1; ++k) {
                    contributor++;
                    const uint32_t g = point_list[k];
                    const R dx = means2D[2 * g] - pfx, dy = means2D[2 * g + 1] - pfy;
                    const R* co = conic_opacity + 4 * (size_t)g;
                    const R power = CF(-0.5) * (co[0] * dx * dx + co[2] * dy * dy) - co[1] * dx * dy;
                    if (power > CF(0.0))
                        continue;
                    const R alpha = RMIN(CF(0.99), co[3] * REXP(power));
                    if (alpha < CF(1.0) / CF(255.0))
                        continue;
                    const R test_T = T * (CF(1.0) - alpha);
                    if (test_T < CF(0.0001)) {
                        h = (h ^ (uint64_t)(0x80000000u | (k - r0))) * 1099511628211ull;
                        break;
                    }
                    for (int ch = 0; ch < 3; ++ch)
                        C[ch] += colors[3 * (size_t)g + ch] * alpha * T;
                    const R w = T * alpha;
                    if (w > weight_max) { weight_max = w; max_id = (int)g; }
                    if (w > pw[g]) pw[g] = w;
                    if (kv_alpha_sum) kv_alpha_sum[k] += w;
                    if (kv_alpha_count) kv_alpha_count[k] += 1;
                    h = (h ^ (uint64_t)(k - r0)) * 1099511628211ull;
                    T = test_T;
                    last_contributor = contributor;
                }
                final_T[pix_id] = T;
                n_contrib[pix_id] = last_contributor;
                for (int ch = 0; ch < 3; ++ch)
                    out_color[(size_t)ch * H * W + pix_id] = C[ch] + T * (R)s->bg[ch];
                point_id[pix_id] = max_id;
                weight_pixel[pix_id] = weight_max;
                if (decision_hash) decision_hash[pix_id] = h;
            }
        }
    }
    for (int t = 0; t < nt; ++t)
        for (int g = 0; g < P; ++g) {
            R v = pw_priv[(size_t)t * P + g];
            if (v > point_weight[g]) point_weight[g] = v;
        }
    free(pw_priv);
}

/* ---- blend backward (backward.cu:399-557) ------------------------------- */
/* One tile's pixels, gradients accumulated into the given arrays at row
 * slot[g] of Gaussian g (at row g when slot is NULL). */
static void SFX(blend_bwd_tile)(const orc_settings* s, int tile, const uint32_t* ranges,
                                const uint32_t* point_list, const R* means2D, const R* colors,
                                const R* conic_opacity, const R* final_T, const uint32_t* n_contrib,
                                const R* dL_dpix, const int32_t* slot, R* dL_dmean2D, R* dL_dconic,
                                R* dL_dopacity, R* dL_dcolors)
{
    const int W = s->image_width, H = s->image_height;
    const int gx = (W + 15) / 16;
    const R ddelx_dx = (R)(0.5 * (double)W), ddely_dy = (R)(0.5 * (double)H);
    const int exact = s->grad_mode == 1;
    {
        const int tx = tile % gx, ty = tile / gx;
        const uint32_t r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
        for (int ly = 0; ly < 16; ++ly) {
            for (int lx = 0; lx < 16; ++lx) {
                const int pxi = tx * 16 + lx, pyi = ty * 16 + ly;
                if (pxi >= W || pyi >= H)
                    continue;
                const int pix_id = W * pyi + pxi;
                const R pfx = (R)pxi, pfy = (R)pyi;
                const R T_final = final_T[pix_id];
                R T = T_final;
                const uint32_t last_contributor = n_contrib[pix_id];
                R dpix[3], accum_rec[3] = { 0, 0, 0 }, last_color[3] = { 0, 0, 0 };
                for (int ch = 0; ch < 3; ++ch) dpix[ch] = dL_dpix[(size_t)ch * H * W + pix_id];
                R last_alpha = 0;
                R bg_dot = 0;
                for (int ch = 0; ch < 3; ++ch) bg_dot += (R)s->bg[ch] * dpix[ch];
                uint32_t contributor = r1 - r0;
                for (uint32_t kk = r1; kk > r0; --kk) {
                    const uint32_t k = kk - 1;
                    contributor--;
                    if (contributor >= last_contributor)
                        continue;
                    const uint32_t g = point_list[k];
                    const size_t o = slot ? (size_t)slot[g] : (size_t)g;
                    const R dx = means2D[2 * g] - pfx, dy = means2D[2 * g + 1] - pfy;
                    const R* co = conic_opacity + 4 * (size_t)g;
                    const R power = CF(-0.5) * (co[0] * dx * dx + co[2] * dy * dy) - co[1] * dx * dy;
                    if (power > CF(0.0))
                        continue;
                    const R G = REXP(power);
                    const R alpha_raw = co[3] * G;
                    const R alpha = RMIN(CF(0.99), alpha_raw);
                    if (alpha < CF(1.0) / CF(255.0))
                        continue;
                    T = T / (CF(1.0) - alpha);
                    const R dchannel_dcolor = alpha * T;
                    R dL_dalpha = 0;
                    for (int ch = 0; ch < 3; ++ch) {
                        const R cc = colors[3 * (size_t)g + ch];
                        accum_rec[ch] = last_alpha * last_color[ch] + (CF(1.0) - last_alpha) * accum_rec[ch];
                        last_color[ch] = cc;
                        dL_dalpha += (cc - accum_rec[ch]) * dpix[ch];
                        dL_dcolors[3 * o + ch] += dchannel_dcolor * dpix[ch];
                    }
                    dL_dalpha *= T;
                    last_alpha = alpha;
                    dL_dalpha += (-T_final / (CF(1.0) - alpha)) * bg_dot;
                    /* exact mode: alpha = min(0.99, o G) has zero slope when clamped */
                    const R dalpha_dog = (exact && alpha_raw > CF(0.99)) ? CF(0.0) : CF(1.0);
                    const R dL_dG = co[3] * dL_dalpha * dalpha_dog;
                    const R gdx = G * dx, gdy = G * dy;
                    const R dG_ddelx = -gdx * co[0] - gdy * co[1];
                    const R dG_ddely = -gdy * co[2] - gdx * co[1];
                    dL_dmean2D[3 * o + 0] += dL_dG * dG_ddelx * ddelx_dx;
                    dL_dmean2D[3 * o + 1] += dL_dG * dG_ddely * ddely_dy;
                    dL_dconic[3 * o + 0] += CF(-0.5) * gdx * dx * dL_dG;
                    dL_dconic[3 * o + 1] += CF(-0.5) * gdx * dy * dL_dG;
                    dL_dconic[3 * o + 2] += CF(-0.5) * gdy * dy * dL_dG;
                    dL_dopacity[o] += G * dL_dalpha * dalpha_dog;
                }
            }
        }
    }
}

void SFX(orc_blend_bwd)(const orc_settings* s, int32_t P, const uint32_t* ranges,
                        const uint32_t* point_list, const R* means2D, const R* colors,
                        const R* conic_opacity, const R* final_T, const uint32_t* n_contrib,
                        const R* dL_dpix, R* dL_dmean2D, R* dL_dconic,
                        R* dL_dopacity, R* dL_dcolors, int32_t nthreads)
{
    const int gx = (s->image_width + 15) / 16, gy = (s->image_height + 15) / 16;
    const int T_ = gx * gy;
    const int nt = nthreads > 0 ? nthreads : 1;
    if (nt == 1 || P <= 0) {
        for (int tile = 0; tile < T_; ++tile)
            SFX(blend_bwd_tile)(s, tile, ranges, point_list, means2D, colors, conic_opacity, final_T,
                                n_contrib, dL_dpix, NULL, dL_dmean2D, dL_dconic, dL_dopacity,
                                dL_dcolors);
        return;
    }
    /* accumulator rows only for the U Gaussians in the list; per thread:
     * mean2D[3U] conic[3U] opacity[U] colour[3U]; tiles dealt out
     * statically so that each run sums in the same order */
    uint32_t K = 0;
    for (int t = 0; t < T_; ++t)
        if (ranges[2 * t + 1] > K) K = ranges[2 * t + 1];
    int32_t* slot = (int32_t*)malloc((size_t)P * sizeof(int32_t));
    for (int32_t g = 0; g < P; ++g) slot[g] = -1;
    int32_t U = 0;
    for (uint32_t k = 0; k < K; ++k)
        if (slot[point_list[k]] < 0) slot[point_list[k]] = U++;
    const size_t per = (size_t)10 * (size_t)(U > 0 ? U : 1);
    R* priv = (R*)calloc((size_t)nt * per, sizeof(R));
    const size_t u3 = 3 * (size_t)U;
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 2) num_threads(nt)
#endif
    for (int tile = 0; tile < T_; ++tile) {
#ifdef _OPENMP
        R* b = priv + (size_t)omp_get_thread_num() * per;
#else
        R* b = priv;
#endif
        SFX(blend_bwd_tile)(s, tile, ranges, point_list, means2D, colors, conic_opacity, final_T,
                            n_contrib, dL_dpix, slot, b, b + u3, b + 2 * u3, b + 2 * u3 + U);
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int g = 0; g < P; ++g) {
        const int32_t u = slot[g];
        if (u < 0)
            continue;
        for (int t = 0; t < nt; ++t) {
            const R* b = priv + (size_t)t * per;
            for (int k = 0; k < 3; ++k) {
                dL_dmean2D[3 * (size_t)g + k] += b[3 * (size_t)u + k];
                dL_dconic[3 * (size_t)g + k] += b[u3 + 3 * (size_t)u + k];
                dL_dcolors[3 * (size_t)g + k] += b[2 * u3 + U + 3 * (size_t)u + k];
            }
            dL_dopacity[g] += b[2 * u3 + u];
        }
    }
    free(priv);
    free(slot);
}

/* ---- preprocess backward (backward.cu:144-396) -------------------------- */
void SFX(orc_preprocess_bwd)(const orc_settings* s, int32_t P, int32_t M, const R* means3D,
                             const R* scales, const R* rotations, const R* opacities,
                             const R* shs, const R* cov3D, const int32_t* radii,
                             const R* dL_dmean2D, const R* dL_dconic, R* dL_dopacity,
                             const R* dL_dcolors, R* dL_dmeans3D, R* dL_dcov3D,
                             R* dL_dsh, R* dL_dscales, R* dL_drot, int32_t nthreads)
{
    const int W = s->image_width, H = s->image_height;
    const float focal_y_f = (float)H / (2.0f * s->tanfovy);
    const float focal_x_f = (float)W / (2.0f * s->tanfovx);
    const R h_x = (R)focal_x_f, h_y = (R)focal_y_f;
    const R tanx = (R)s->tanfovx, tany = (R)s->tanfovy;
    const R mod = (R)s->scale_modifier;
    const int exact = s->grad_mode == 1;
    const int deg = s->sh_degree;
    const float* V = s->viewmatrix;
    const float* Pm = s->projmatrix;
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 256) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = 0; i < P; ++i) {
        if (!(radii[i] > 0))
            continue;
        const R mx = means3D[3 * i], my = means3D[3 * i + 1], mz = means3D[3 * i + 2];
        const R* c3 = cov3D + 6 * (size_t)i;
        SFX(cov2d_t) cv;
        SFX(cov2d)(V, mx, my, mz, h_x, h_y, tanx, tany, c3, &cv);
        const R x_grad_mul = (cv.txtz < -cv.limx || cv.txtz > cv.limx) ? CF(0.0) : CF(1.0);
        const R y_grad_mul = (cv.tytz < -cv.limy || cv.tytz > cv.limy) ? CF(0.0) : CF(1.0);
        R a = cv.a, b = cv.b, c = cv.c;
        R mask_a = CF(1.0), mask_c = CF(1.0);
        if (!exact || s->variant == 1) {
            /* backward.cu:197-199 always adds DILATE_PIXEL (double literal) */
            a = (R)((double)a + 0.3);
            c = (R)((double)c + 0.3);
        } else if (s->use_filter) {
            mask_a = a > CF(0.3) ? CF(1.0) : CF(0.0);
            mask_c = c > CF(0.3) ? CF(1.0) : CF(0.0);
            a = RMAX(a, CF(0.3));
            c = RMAX(c, CF(0.3));
        }
        const R dcx = dL_dconic[3 * (size_t)i + 0];
        const R dcy = dL_dconic[3 * (size_t)i + 1];
        const R dcz = dL_dconic[3 * (size_t)i + 2];
        const R denom = a * c - b * b;
        R dL_da = 0, dL_db = 0, dL_dc = 0;
        const R denom2inv = exact ? CF(1.0) / (denom * denom)
                                  : CF(1.0) / ((denom * denom) + CF(0.0000001));
        R* dcov = dL_dcov3D + 6 * (size_t)i;
        const R (*T)[3] = cv.T;
        if (denom2inv != CF(0.0)) {
            dL_da = denom2inv * (-c * c * dcx + CF(2.0) * b * c * dcy + (denom - a * c) * dcz);
            dL_dc = denom2inv * (-a * a * dcz + CF(2.0) * a * b * dcy + (denom - a * c) * dcx);
            dL_db = denom2inv * CF(2.0) * (b * c * dcx - (denom + CF(2.0) * b * b) * dcy + a * b * dcz);
            dL_da *= mask_a;
            dL_dc *= mask_c;
            dcov[0] = (T[0][0] * T[0][0] * dL_da + T[0][0] * T[1][0] * dL_db + T[1][0] * T[1][0] * dL_dc);
            dcov[3] = (T[0][1] * T[0][1] * dL_da + T[0][1] * T[1][1] * dL_db + T[1][1] * T[1][1] * dL_dc);
            dcov[5] = (T[0][2] * T[0][2] * dL_da + T[0][2] * T[1][2] * dL_db + T[1][2] * T[1][2] * dL_dc);
            dcov[1] = CF(2.0) * T[0][0] * T[0][1] * dL_da + (T[0][0] * T[1][1] + T[0][1] * T[1][0]) * dL_db + CF(2.0) * T[1][0] * T[1][1] * dL_dc;
            dcov[2] = CF(2.0) * T[0][0] * T[0][2] * dL_da + (T[0][0] * T[1][2] + T[0][2] * T[1][0]) * dL_db + CF(2.0) * T[1][0] * T[1][2] * dL_dc;
            dcov[4] = CF(2.0) * T[0][2] * T[0][1] * dL_da + (T[0][1] * T[1][2] + T[0][2] * T[1][1]) * dL_db + CF(2.0) * T[1][1] * T[1][2] * dL_dc;
        } else {
            for (int k = 0; k < 6; ++k) dcov[k] = 0;
        }
        /* Vrk[row][col] (symmetric) */
        const R Vk[3][3] = { { c3[0], c3[1], c3[2] }, { c3[1], c3[3], c3[4] }, { c3[2], c3[4], c3[5] } };
        const R dL_dT00 = CF(2.0) * (T[0][0] * Vk[0][0] + T[0][1] * Vk[0][1] + T[0][2] * Vk[0][2]) * dL_da +
                          (T[1][0] * Vk[0][0] + T[1][1] * Vk[0][1] + T[1][2] * Vk[0][2]) * dL_db;
        const R dL_dT01 = CF(2.0) * (T[0][0] * Vk[1][0] + T[0][1] * Vk[1][1] + T[0][2] * Vk[1][2]) * dL_da +
                          (T[1][0] * Vk[1][0] + T[1][1] * Vk[1][1] + T[1][2] * Vk[1][2]) * dL_db;
        const R dL_dT02 = CF(2.0) * (T[0][0] * Vk[2][0] + T[0][1] * Vk[2][1] + T[0][2] * Vk[2][2]) * dL_da +
                          (T[1][0] * Vk[2][0] + T[1][1] * Vk[2][1] + T[1][2] * Vk[2][2]) * dL_db;
        const R dL_dT10 = CF(2.0) * (T[1][0] * Vk[0][0] + T[1][1] * Vk[0][1] + T[1][2] * Vk[0][2]) * dL_dc +
                          (T[0][0] * Vk[0][0] + T[0][1] * Vk[0][1] + T[0][2] * Vk[0][2]) * dL_db;
        const R dL_dT11 = CF(2.0) * (T[1][0] * Vk[1][0] + T[1][1] * Vk[1][1] + T[1][2] * Vk[1][2]) * dL_dc +
                          (T[0][0] * Vk[1][0] + T[0][1] * Vk[1][1] + T[0][2] * Vk[1][2]) * dL_db;
        const R dL_dT12 = CF(2.0) * (T[1][0] * Vk[2][0] + T[1][1] * Vk[2][1] + T[1][2] * Vk[2][2]) * dL_dc +
                          (T[0][0] * Vk[2][0] + T[0][1] * Vk[2][1] + T[0][2] * Vk[2][2]) * dL_db;
        const R (*Wm)[3] = cv.W;
        const R dL_dJ00 = Wm[0][0] * dL_dT00 + Wm[0][1] * dL_dT01 + Wm[0][2] * dL_dT02;
        const R dL_dJ02 = Wm[2][0] * dL_dT00 + Wm[2][1] * dL_dT01 + Wm[2][2] * dL_dT02;
        const R dL_dJ11 = Wm[1][0] * dL_dT10 + Wm[1][1] * dL_dT11 + Wm[1][2] * dL_dT12;
        const R dL_dJ12 = Wm[2][0] * dL_dT10 + Wm[2][1] * dL_dT11 + Wm[2][2] * dL_dT12;
        const R tz = CF(1.0) / cv.t[2];
        const R tz2 = tz * tz, tz3 = tz2 * tz;
        const R dL_dtx = x_grad_mul * -h_x * tz2 * dL_dJ02;
        const R dL_dty = y_grad_mul * -h_y * tz2 * dL_dJ12;
        const R dL_dtz = -h_x * tz2 * dL_dJ00 - h_y * tz2 * dL_dJ11 + (CF(2.0) * h_x * cv.t[0]) * tz3 * dL_dJ02 +
                         (CF(2.0) * h_y * cv.t[1]) * tz3 * dL_dJ12;
        R dm[3];
        dm[0] = (R)V[0] * dL_dtx + (R)V[1] * dL_dty + (R)V[2] * dL_dtz;
        dm[1] = (R)V[4] * dL_dtx + (R)V[5] * dL_dty + (R)V[6] * dL_dtz;
        dm[2] = (R)V[8] * dL_dtx + (R)V[9] * dL_dty + (R)V[10] * dL_dtz;

        /* projection of the mean (backward.cu:370-387) */
        R hom[4];
        SFX(xf4)(Pm, mx, my, mz, hom);
        const R m_w = CF(1.0) / (hom[3] + CF(0.0000001));
        const R mul1 = hom[0] * m_w * m_w;
        const R mul2 = hom[1] * m_w * m_w;
        const R g2x = dL_dmean2D[3 * (size_t)i], g2y = dL_dmean2D[3 * (size_t)i + 1];
        dm[0] += ((R)Pm[0] * m_w - (R)Pm[3] * mul1) * g2x + ((R)Pm[1] * m_w - (R)Pm[3] * mul2) * g2y;
        dm[1] += ((R)Pm[4] * m_w - (R)Pm[7] * mul1) * g2x + ((R)Pm[5] * m_w - (R)Pm[7] * mul2) * g2y;
        dm[2] += ((R)Pm[8] * m_w - (R)Pm[11] * mul1) * g2x + ((R)Pm[9] * m_w - (R)Pm[11] * mul2) * g2y;

        /* SH backward (backward.cu:20-139) */
        if (shs) {
            const R* sh = shs + (size_t)i * M * 3;
            R* dsh = dL_dsh + (size_t)i * M * 3;
            R d[3], dor[3];
            SFX(view_dir)(s->campos, mx, my, mz, d, dor);
            const R x = d[0], y = d[1], z = d[2];
            R dRGB[3];
            for (int ch = 0; ch < 3; ++ch) {
                R v = SFX(sh_eval)(deg, sh, ch, x, y, z) + CF(0.5);
                dRGB[ch] = (v < CF(0.0)) ? CF(0.0) : dL_dcolors[3 * (size_t)i + ch];
            }
            R ddx[3] = { 0, 0, 0 }, ddy[3] = { 0, 0, 0 }, ddz[3] = { 0, 0, 0 };
            for (int ch = 0; ch < 3; ++ch) {
#define SHv(k) sh[(k) * 3 + ch]
#define DSH(k) dsh[(k) * 3 + ch]
                const R g = dRGB[ch];
                DSH(0) = SFX(SHC0) * g;
                if (deg > 0) {
                    const R C1 = SFX(SHC1);
                    DSH(1) = -C1 * y * g;
                    DSH(2) = C1 * z * g;
                    DSH(3) = -C1 * x * g;
                    ddx[ch] = -C1 * SHv(3);
                    ddy[ch] = -C1 * SHv(1);
                    ddz[ch] = C1 * SHv(2);
                    if (deg > 1) {
                        const R* C2 = SFX(SHC2);
                        const R xx = x * x, yy = y * y, zz = z * z;
                        const R xy = x * y, yz = y * z, xz = x * z;
                        DSH(4) = C2[0] * xy * g;
                        DSH(5) = C2[1] * yz * g;
                        DSH(6) = C2[2] * (CF(2.0) * zz - xx - yy) * g;
                        DSH(7) = C2[3] * xz * g;
                        DSH(8) = C2[4] * (xx - yy) * g;
                        ddx[ch] += C2[0] * y * SHv(4) + C2[2] * CF(2.0) * -x * SHv(6) + C2[3] * z * SHv(7) + C2[4] * CF(2.0) * x * SHv(8);
                        ddy[ch] += C2[0] * x * SHv(4) + C2[1] * z * SHv(5) + C2[2] * CF(2.0) * -y * SHv(6) + C2[4] * CF(2.0) * -y * SHv(8);
                        ddz[ch] += C2[1] * y * SHv(5) + C2[2] * CF(2.0) * CF(2.0) * z * SHv(6) + C2[3] * x * SHv(7);
                        if (deg > 2) {
                            const R* C3 = SFX(SHC3);
                            DSH(9) = C3[0] * y * (CF(3.0) * xx - yy) * g;
                            DSH(10) = C3[1] * xy * z * g;
                            DSH(11) = C3[2] * y * (CF(4.0) * zz - xx - yy) * g;
                            DSH(12) = C3[3] * z * (CF(2.0) * zz - CF(3.0) * xx - CF(3.0) * yy) * g;
                            DSH(13) = C3[4] * x * (CF(4.0) * zz - xx - yy) * g;
                            DSH(14) = C3[5] * z * (xx - yy) * g;
                            DSH(15) = C3[6] * x * (xx - CF(3.0) * yy) * g;
                            ddx[ch] += (C3[0] * SHv(9) * CF(3.0) * CF(2.0) * xy + C3[1] * SHv(10) * yz +
                                        C3[2] * SHv(11) * CF(-2.0) * xy + C3[3] * SHv(12) * CF(-3.0) * CF(2.0) * xz +
                                        C3[4] * SHv(13) * (CF(-3.0) * xx + CF(4.0) * zz - yy) +
                                        C3[5] * SHv(14) * CF(2.0) * xz + C3[6] * SHv(15) * CF(3.0) * (xx - yy));
                            ddy[ch] += (C3[0] * SHv(9) * CF(3.0) * (xx - yy) + C3[1] * SHv(10) * xz +
                                        C3[2] * SHv(11) * (CF(-3.0) * yy + CF(4.0) * zz - xx) +
                                        C3[3] * SHv(12) * CF(-3.0) * CF(2.0) * yz + C3[4] * SHv(13) * CF(-2.0) * xy +
                                        C3[5] * SHv(14) * CF(-2.0) * yz + C3[6] * SHv(15) * CF(-3.0) * CF(2.0) * xy);
                            ddz[ch] += (C3[1] * SHv(10) * xy + C3[2] * SHv(11) * CF(4.0) * CF(2.0) * yz +
                                        C3[3] * SHv(12) * CF(3.0) * (CF(2.0) * zz - xx - yy) +
                                        C3[4] * SHv(13) * CF(4.0) * CF(2.0) * xz + C3[5] * SHv(14) * (xx - yy));
                        }
                    }
                }
#undef SHv
#undef DSH
            }
            const R dL_ddir[3] = { ddx[0] * dRGB[0] + ddx[1] * dRGB[1] + ddx[2] * dRGB[2],
                                   ddy[0] * dRGB[0] + ddy[1] * dRGB[1] + ddy[2] * dRGB[2],
                                   ddz[0] * dRGB[0] + ddz[1] * dRGB[1] + ddz[2] * dRGB[2] };
            /* dnormvdv (auxiliary.h:107-117) */
            const R sum2 = dor[0] * dor[0] + dor[1] * dor[1] + dor[2] * dor[2];
            const R invsum32 = CF(1.0) / RSQRT(sum2 * sum2 * sum2);
            dm[0] += ((sum2 - dor[0] * dor[0]) * dL_ddir[0] - dor[1] * dor[0] * dL_ddir[1] - dor[2] * dor[0] * dL_ddir[2]) * invsum32;
            dm[1] += (-dor[0] * dor[1] * dL_ddir[0] + (sum2 - dor[1] * dor[1]) * dL_ddir[1] - dor[2] * dor[1] * dL_ddir[2]) * invsum32;
            dm[2] += (-dor[0] * dor[2] * dL_ddir[0] - dor[1] * dor[2] * dL_ddir[1] + (sum2 - dor[2] * dor[2]) * dL_ddir[2]) * invsum32;
        }
        for (int k = 0; k < 3; ++k) dL_dmeans3D[3 * (size_t)i + k] = dm[k];

        /* cov3D backward (backward.cu:278-341) */
        if (scales) {
            const R r = rotations[4 * i], x = rotations[4 * i + 1], y = rotations[4 * i + 2], z = rotations[4 * i + 3];
            R Rm[3][3];
            SFX(quat_rot)(r, x, y, z, Rm);
            const R sv[3] = { mod * scales[3 * i], mod * scales[3 * i + 1], mod * scales[3 * i + 2] };
            R Mm[3][3];
            for (int j = 0; j < 3; ++j)
                for (int q = 0; q < 3; ++q) Mm[j][q] = sv[q] * Rm[j][q];
            const R* dc = dcov;
            /* dL_dSigma (glm columns) */
            const R dS[3][3] = { { dc[0], CF(0.5) * dc[1], CF(0.5) * dc[2] },
                                 { CF(0.5) * dc[1], dc[3], CF(0.5) * dc[4] },
                                 { CF(0.5) * dc[2], CF(0.5) * dc[4], dc[5] } };
            /* dL_dM = (2 M) * dL_dSigma */
            R dM[3][3];
            for (int j = 0; j < 3; ++j)
                for (int q = 0; q < 3; ++q)
                    dM[j][q] = (CF(2.0) * Mm[0][q]) * dS[j][0] + (CF(2.0) * Mm[1][q]) * dS[j][1] + (CF(2.0) * Mm[2][q]) * dS[j][2];
            /* Rt = transpose(R), dMt = transpose(dM) */
            R ds[3];
            for (int q = 0; q < 3; ++q)
                ds[q] = Rm[0][q] * dM[0][q] + Rm[1][q] * dM[1][q] + Rm[2][q] * dM[2][q];
            R dMt[3][3];
            for (int j = 0; j < 3; ++j)
                for (int q = 0; q < 3; ++q) dMt[j][q] = dM[q][j];
            for (int q = 0; q < 3; ++q) {
                dMt[0][q] *= sv[0];
                dMt[1][q] *= sv[1];
                dMt[2][q] *= sv[2];
            }
            const R smul = exact ? mod : CF(1.0);
            for (int q = 0; q < 3; ++q) dL_dscales[3 * (size_t)i + q] = ds[q] * smul;
            R* dq = dL_drot + 4 * (size_t)i;
            dq[0] = CF(2.0) * z * (dMt[0][1] - dMt[1][0]) + CF(2.0) * y * (dMt[2][0] - dMt[0][2]) + CF(2.0) * x * (dMt[1][2] - dMt[2][1]);
            dq[1] = CF(2.0) * y * (dMt[1][0] + dMt[0][1]) + CF(2.0) * z * (dMt[2][0] + dMt[0][2]) + CF(2.0) * r * (dMt[1][2] - dMt[2][1]) - CF(4.0) * x * (dMt[2][2] + dMt[1][1]);
            dq[2] = CF(2.0) * x * (dMt[1][0] + dMt[0][1]) + CF(2.0) * r * (dMt[2][0] - dMt[0][2]) + CF(2.0) * z * (dMt[1][2] + dMt[2][1]) - CF(4.0) * y * (dMt[2][2] + dMt[0][0]);
            dq[3] = CF(2.0) * r * (dMt[0][1] - dMt[1][0]) + CF(2.0) * x * (dMt[2][0] + dMt[0][2]) + CF(2.0) * y * (dMt[1][2] + dMt[2][1]) - CF(4.0) * z * (dMt[1][1] + dMt[0][0]);
        }
        if (exact && s->variant == 0 && opacities && opacities[i] < CF(1e-6))
            dL_dopacity[i] = 0;   /* o' = max(o, 1e-6) has zero slope below the floor */
    }
}

#undef CF
