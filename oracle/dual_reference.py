"""FP64 forward-mode (dual-number) reference of the step and its derivatives.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the adjoint,
never by the product (fibrenet_amd/), which runs on the GPU or not at all.

It restates the fast path of the reference in numpy on dual numbers: every
state array carries its value and its derivatives along M input directions
(force scales of springs, angles and contacts, initial positions, initial
velocity_old).  Running the step on these is forward-mode differentiation,
independent of the hand-derived adjoint equations it checks.  Each function
follows the reference's own formula:
  drift, damping, gravity, walls      kernel.cl:195-265 (AdvancePosition)
  springs (HR, HR_FL with failure)    kernel.cl:366-441 (ResolveSprings)
  angles (acos form, clamp)           kernel.cl:268-363 (ResolveAngles)
  contacts lj / hs / st               kernel.cl:792-1011 (ResolveContacts)
  kick                                kernel.cl:444-464 (AdvanceVelocity)
  contact lists                       kernel.cl:487-506, 532-789 (FillGrid,
                                      BuildContactList) on the grid of
                                      SimulationManager.cpp:2177-2301
  rebuild cadence                     SimulationManager.cpp:151-181
Lengths under native_sqrt take the float-rounded value (as the reference
computes them) with the exact derivative; `round_lengths=False` uses exact
square roots throughout (a smooth forward, for the central-difference pin).
Contact lists and bond statuses are the reference's rules at this run's own
trajectory and are constant in the derivative.  Arithmetic is fp64 without
fused multiply-adds, so values agree with the engine and the C oracle to
rounding, not bit for bit.

Pins (tests/test_oracle.py): values against the C oracle's trajectory;
tangents against central differences of this same forward with exact
square roots.
"""
import numpy as np


class Dual:
    """v: values (any shape); d: derivatives, shape (M,) + v.shape."""
    __slots__ = ("v", "d")
    __array_ufunc__ = None      # ndarray (op) Dual defers to Dual's reflected operators

    def __init__(self, v, d):
        self.v = v
        self.d = d

    @staticmethod
    def const(v, m):
        v = np.asarray(v, np.float64)
        return Dual(v, np.zeros((m,) + v.shape))

    def __add__(a, b):
        if isinstance(b, Dual):
            return Dual(a.v + b.v, a.d + b.d)
        return Dual(a.v + b, a.d)

    __radd__ = __add__

    def __sub__(a, b):
        if isinstance(b, Dual):
            return Dual(a.v - b.v, a.d - b.d)
        return Dual(a.v - b, a.d)

    def __rsub__(a, b):
        return Dual(b - a.v, -a.d)

    def __neg__(a):
        return Dual(-a.v, -a.d)

    def __mul__(a, b):
        if isinstance(b, Dual):
            return Dual(a.v * b.v, a.d * b.v + a.v * b.d)
        return Dual(a.v * b, a.d * b)

    __rmul__ = __mul__

    def __truediv__(a, b):
        if isinstance(b, Dual):
            q = a.v / b.v
            return Dual(q, (a.d - q * b.d) / b.v)
        return Dual(a.v / b, a.d / b)

    def __rtruediv__(a, b):
        q = b / a.v
        return Dual(q, -q * a.d / a.v)

    def take(self, idx):
        return Dual(self.v[idx], self.d[:, idx])

    def where(self, mask, other=0.0):
        """Elementwise: self where mask, else the constant `other`."""
        return Dual(np.where(mask, self.v, other), np.where(mask, self.d, 0.0))


def dsqrt(a):
    r = np.sqrt(a.v)
    safe = np.where(r > 0.0, r, 1.0)
    return Dual(r, np.where(r > 0.0, a.d / (2.0 * safe), 0.0))


def native_sqrt(a, rounded=True):
    """sqrt as the reference's native_sqrt(double): the float-rounded value
    (kernel.cl under clang OpenCL), differentiated as the exact root."""
    e = dsqrt(a)
    if rounded:
        e.v = np.sqrt(a.v.astype(np.float32)).astype(np.float64)   # sqrtf((float)x)
    return e


def darccos(a):
    return Dual(np.arccos(a.v), -a.d / np.sqrt(np.maximum(1.0 - a.v * a.v, 1e-300)))


def scatter_add(target, rows, val):
    n = len(target.v)
    target.v = target.v + np.bincount(rows, weights=val.v, minlength=n)
    target.d = target.d + np.stack([np.bincount(rows, weights=val.d[m], minlength=n) for m in range(len(val.d))])


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


DIRECTIONS = ("spring_scale", "angle_scale", "contact_scale", "position0", "velocity_old0")


class DualRun:
    """A fast-path run of one network from its records (generate.Network
    fields) with derivatives along DIRECTIONS: the three force scales (all
    springs, all angles, every contact pair type, multiplied by 1 + eps)
    and the initial positions / velocity_old moved along u_pos / u_vel."""

    def __init__(self, beads, info, springs, angles, pairs, settings, u_pos, u_vel, round_lengths=True,
                 scales=(1.0, 1.0, 1.0)):
        m = len(DIRECTIONS)
        self.M = m
        self.rounded = round_lengths
        s = np.asarray(settings).reshape(())
        self.s = s
        self.N = N = len(beads)
        self.info = info
        self.springs = springs.copy()
        self.angles = angles.copy()
        self.pairs = pairs
        self.scales = [float(x) for x in scales]

        def vec(field):
            return [np.array(beads[field][c], np.float64) for c in "xyz"]

        pos, img, w, f = vec("position"), vec("image_position"), vec("velocity_old"), vec("force")
        up = np.asarray(u_pos, np.float64).reshape(N, 3)
        uv = np.asarray(u_vel, np.float64).reshape(N, 3)

        def with_dir(v, k, u):
            d = np.zeros((m, N))
            if u is not None:
                d[k] = u
            return Dual(v, d)

        self.pos = [with_dir(pos[c], 3, up[:, c]) for c in range(3)]
        self.img = [with_dir(img[c], 3, up[:, c]) for c in range(3)]
        self.w = [with_dir(w[c], 4, uv[:, c]) for c in range(3)]
        self.f = [with_dir(f[c], 0, None) for c in range(3)]
        self.sp_status = np.array(springs["status"], np.int64)
        self.an_status = np.array(angles["status"], np.int64)
        self.lam = np.array(info["lambda_dt_mass_inv"], np.float64)
        self.hlf = np.array(info["half_dt_mass_inv"], np.float64)
        self.hlf2 = np.array(info["half_dt2_mass_inv"], np.float64)
        self.rad = np.array(info["radius"], np.float64)
        self.list_i = self.list_j = self.list_t = np.zeros(0, np.int64)

    # -- scales as dual numbers ------------------------------------------------------------
    def _scale(self, k, n):
        d = np.zeros((self.M, n))
        d[k] = self.scales[k]
        return Dual(np.full(n, self.scales[k]), d)

    # -- contact list (FillGrid + BuildContactList, two-sided) ------------------------------------
    def rebuild(self):
        s, N = self.s, self.N
        px = np.stack([p.v for p in self.pos], axis=1)
        n = np.array([s["cells_x"], s["cells_y"], s["cells_z"]], np.int64)
        step = np.array([s["grid_step"][c] for c in "xyz"], np.float64)
        cell = np.floor(px / step).astype(np.int64)
        hashed = np.all((cell >= 0) & (cell < n), axis=1)
        idx = np.nonzero(hashed)[0]
        ci = cell[idx]
        i, j = self._candidates(px[idx])
        dc = np.abs(ci[i] - ci[j])
        if int(s["bc_type"]) == 0:
            near = np.all((dc <= 1) | (dc == n - 1), axis=1)
        else:
            near = np.all(dc <= 1, axis=1)
        i, j = idx[i[near]], idx[j[near]]
        tag, loc = np.asarray(self.info["tag"]), np.asarray(self.info["loc_id"])
        coll = (tag[i] != tag[j]) | (np.abs(loc[i] - loc[j]) > 1)
        i, j = i[coll], j[coll]
        r = px[i] - px[j]
        len2 = np.sum(r * r, axis=1)
        if int(s["bc_type"]) == 0:
            hbs = float(s["half_box_sq"])
            L = np.array([s["box_size"][c] for c in "xyz"], np.float64)
            far = len2 > hbs
            sh = (r * r > hbs) & far[:, None]
            r = np.where(sh, r + np.where(r > 0, -L, L), r)
            len2 = np.sum(r * r, axis=1)
        T = int(s["types_num"])
        ty = np.asarray(self.info["type"], np.int64)
        lo, hi = np.minimum(ty[i], ty[j]), np.maximum(ty[i], ty[j])
        t = T * lo + hi - lo * (lo + 1) // 2
        ptype = np.asarray(self.pairs["type"])[t]
        skin = np.asarray(self.info["skin_radius"], np.float64)
        reach2 = np.where(ptype == 2, (np.maximum(self.rad[i], self.rad[j]) + np.minimum(skin[i], skin[j])) ** 2,
                          float(s["cutoff_sq"]))
        keep = len2 < reach2
        i, j, t = i[keep], j[keep], t[keep]
        self.list_i = np.concatenate([i, j])
        self.list_j = np.concatenate([j, i])
        self.list_t = np.concatenate([t, t])

    def _candidates(self, p):
        """Index pairs (i < j) of p that may be listed: every pair for a few
        beads, else every pair within the largest listing reach (a k-d tree
        query; the listing rule is applied after)."""
        if len(p) <= 2000:
            return np.triu_indices(len(p), 1)
        from scipy.spatial import cKDTree
        s = self.s
        skin = np.asarray(self.info["skin_radius"], np.float64)
        reach = max(float(np.sqrt(s["cutoff_sq"])), float(self.rad.max() + skin.max()))
        if int(s["bc_type"]) == 0:
            L = np.array([s["box_size"][c] for c in "xyz"], np.float64)
            pairs = cKDTree(np.mod(p, L), boxsize=L).query_pairs(reach * (1 + 1e-9), output_type="ndarray")
        else:
            pairs = cKDTree(p).query_pairs(reach * (1 + 1e-9), output_type="ndarray")
        pairs = np.sort(pairs, axis=1)
        return pairs[:, 0], pairs[:, 1]

    # -- one step ----------------------------------------------------------------------------------
    def step(self):
        s = self.s
        dt = float(s["dt"])
        lam, h, h2 = self.lam, self.hlf, self.hlf2
        # AdvancePosition (kernel.cl:195-265)
        v = [self.w[c] + self.f[c] * lam for c in range(3)]
        delta = [self.w[c] * dt + self.f[c] * h2 for c in range(3)]
        self.pos = [self.pos[c] + delta[c] for c in range(3)]
        self.img = [self.img[c] + delta[c] for c in range(3)]
        f_old = self.f
        g = [float(s["gravity"][c]) for c in "xyz"]
        damp = float(s["damping"])
        f = [v[c] * damp + g[c] for c in range(3)]
        L = [float(s["box_size"][c]) for c in "xyz"]
        if int(s["bc_type"]) == 0:
            for c in (2, 1, 0):
                p = self.pos[c]
                p.v = np.where(p.v < 0.0, L[c], p.v)
            for c in (2, 1, 0):
                p = self.pos[c]
                p.v = np.where(p.v > L[c], 0.0, p.v)
        else:
            kw = float(s["wall_contact_stiffness"])
            for c in range(3):
                p = self.pos[c]
                lo = p.v < 0.0
                hi = p.v > L[c]
                f[c] = f[c] + (p * (-kw)).where(lo) + ((p - L[c]) * (-kw)).where(hi)
        return v, f, f_old

    def bonds_and_contacts(self, f, step_index):
        self._springs(f, step_index)
        self._angles(f)
        self._contacts(f)

    def _springs(self, f, step_index):
        sp = self.springs
        if len(sp) == 0:
            return
        p1, p2 = np.asarray(sp["p1"]), np.asarray(sp["p2"])
        d = [self.img[c].take(p1) - self.img[c].take(p2) for c in range(3)]
        ln = native_sqrt(_dot(d, d), self.rounded)
        factor = np.asarray(sp["rest_length"], np.float64) - ln
        inv = (1.0 / ln.where(ln.v != 0.0, 1.0)).where(ln.v != 0.0)
        ty = np.asarray(sp["type"])
        live = (self.sp_status == 1) & ((ty == 0) | (ty == 1))
        stress = -np.asarray(sp["c1"], np.float64) * factor.v * inv.v
        broke = live & (ty == 1) & (np.abs(stress) > np.asarray(sp["c2"], np.float64))
        self.sp_status = np.where(broke, 0, self.sp_status)
        live = live & ~broke
        k = self._scale(0, len(sp)) * np.asarray(sp["stiffness"], np.float64)
        mag = factor * k
        F = [(d[c] * inv * mag).where(live) for c in range(3)]
        for c in range(3):
            scatter_add(f[c], p1, F[c])
            scatter_add(f[c], p2, -F[c])

    def _angles(self, f):
        an = self.angles
        if len(an) == 0:
            return
        ty = np.asarray(an["type"])
        s1, s2 = np.asarray(an["s1"]), np.asarray(an["s2"])
        dead = (self.an_status == 1) & (ty == 1) & ((self.sp_status[s1] == 0) | (self.sp_status[s2] == 0))
        self.an_status = np.where(dead, 0, self.an_status)
        live = self.an_status == 1
        b1, b2, b3 = np.asarray(an["b1"]), np.asarray(an["b2"]), np.asarray(an["b3"])
        ba = [self.img[c].take(b1) - self.img[c].take(b2) for c in range(3)]
        bc = [self.img[c].take(b3) - self.img[c].take(b2) for c in range(3)]
        l_ba, l_bc = dsqrt(_dot(ba, ba)), dsqrt(_dot(bc, bc))
        dotp = _dot(ba, bc) / (l_ba * l_bc)
        clamp_lo, clamp_hi = dotp.v < -1.0, dotp.v > 1.0
        dotp = dotp.where(~clamp_lo, -0.999999).where(~clamp_hi, 0.999999)
        ang = darccos(dotp)
        k = self._scale(1, len(an)) * np.asarray(an["c1"], np.float64)
        mag = (ang - np.asarray(an["angle"], np.float64)) * k * -2.0
        fa, fc = mag / l_ba, mag / l_bc
        cb = [-x for x in bc]
        cr = _cross(ba, bc)
        pa = _cross(ba, cr)
        pc = _cross(cb, cr)
        out = []
        for p, fm in ((pa, fa), (pc, fc)):
            ln = dsqrt(_dot(p, p))
            ok = ln.v != 0.0
            scale = (fm / ln.where(ok, 1.0)).where(ok)
            out.append([(p[c] * scale).where(live) for c in range(3)])
        pa, pc = out
        for c in range(3):
            scatter_add(f[c], b1, pa[c])
            scatter_add(f[c], b3, pc[c])
            scatter_add(f[c], b2, -(pa[c] + pc[c]))

    def _contacts(self, f):
        i, j, t = self.list_i, self.list_j, self.list_t
        if len(i) == 0:
            return
        s = self.s
        r = [self.pos[c].take(i) - self.pos[c].take(j) for c in range(3)]
        len2 = _dot(r, r)
        hbs = float(s["half_box_sq"])
        L = [float(s["box_size"][c]) for c in "xyz"]
        far = len2.v > hbs
        for c in range(3):   # minimum image, ResolveContacts (kernel.cl:844-870)
            sh = far & (r[c].v * r[c].v > hbs)
            r[c] = r[c] + np.where(sh, np.where(r[c].v > 0, -L[c], L[c]), 0.0)
        len2 = _dot(r, r)
        pr = self.pairs[t]
        ptype = np.asarray(pr["type"])
        c1, c2, c3 = (np.asarray(pr[k], np.float64) for k in ("c1", "c2", "c3"))
        c4, c5, c6 = (np.asarray(pr[k], np.float64) for k in ("c4", "c5", "c6"))
        rr = native_sqrt(len2, self.rounded)
        sig = self._scale(2, len(i))
        # lj (kernel.cl:875-919): |F|/r along r
        fr2 = (c1 * c1) / len2
        fr6 = fr2 * fr2 * fr2
        lj = (fr6 * (fr6 - 0.5) * (48.0 * c2)) / len2 - c4
        on_lj = (ptype == 1) & (len2.v < float(s["cutoff_sq"])) & (rr.v < c3)
        # hs (kernel.cl:922-957)
        pen = (self.rad[i] + self.rad[j]) - rr
        on_hs = (ptype == 2) & (pen.v > 0.0)
        hs = pen * c1
        # st (kernel.cl:960-999)
        st = (c2 - rr) * c1 + (c6 - rr) * c5
        on_st = (ptype == 3) & (len2.v < c4)
        coef = lj.where(on_lj) + hs.where(on_hs) + st.where(on_st)
        on = on_lj | on_hs | on_st
        inv = (1.0 / rr.where(rr.v != 0.0, 1.0)).where(rr.v != 0.0)
        coef = coef * sig
        for c in range(3):
            scatter_add(f[c], i, (coef * (r[c] * inv)).where(on))

    def run(self, n_steps, grid_update_freq, group, sample_steps):
        """n_steps steps; returns the dual samples (sum over `group` of the
        x-force) at sample_steps, values and derivatives (M, n_samples)."""
        group = np.asarray(group) != 0
        want = {int(k): q for q, k in enumerate(sample_steps)}
        vals = np.zeros(len(want))
        ders = np.zeros((self.M, len(want)))
        for n in range(n_steps):
            v, f, f_old = self.step()
            if n % grid_update_freq == 0:
                self.rebuild()
            self.bonds_and_contacts(f, n)
            h = self.hlf
            self.w = [self.w[c] + (f[c] + f_old[c]) * h for c in range(3)]
            self.f = f
            if n in want:
                q = want[n]
                vals[q] = float(np.sum(f[0].v[group]))
                ders[:, q] = np.sum(f[0].d[:, group], axis=1)
        return vals, ders

    def state(self):
        """Values of positions, images, velocity_old and force, (N, 3) each."""
        st = lambda a: np.stack([x.v for x in a], axis=1)
        return {"position": st(self.pos), "image_position": st(self.img), "velocity_old": st(self.w),
                "force": st(self.f)}
