"""Writes the golden fixtures under tests/golden/ (run from the repo root).

Every expected value here is computed independently of the oracle:
* fft_vectors.npz       numpy.fft of seeded random inputs (n = 8 ... 384, incl. a prime)
* gpe2d_matlab.npz      numpy transcription of the imaginary-time loop of GPUE's
                        matlab/GPE_2d.m (lines 87-110: GVr*exp(-|psi|^2 U Dt/2hbar)*psi,
                        ifft2(GKp*fft2(psi)), repeat, renormalise), no rotation (l = 0),
                        on GPUE's grid (src/operators.cu:336-351)
* gpue_gauge_2d.npz     numpy transcription of GPUE's step order with the per-axis
                        gauge operators (src/evolution.cu:163-257, 727-870) and
                        V += 1/2 m |A|^2 (src/operators.cu:1006-1007), imaginary and real time
* gpue_gauge_3d.npz     the same in 3D (src/evolution.cu:5-160)
* gpue_gauge_box.npz    the same on a 1D line (only the x operator; the y grid of an
                        unused axis is 0, so A_x = 0 and V gains 1/2 (x omega)^2 from A_y),
                        a rectangular 2D box (n_x != n_y, xMax != yMax, gammaY != 1) and a
                        3D box (n_x != n_y = n_z, three different extents)
* vortex_cases.npz      phase fields with vortices imprinted inside known plaquettes and
                        the markers GPUE's skip/sign/mask rules give for them, derived
                        from the imprint geometry (src/tracker.cu:102-161)
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def gpue_grid(n, xmax):
    # x_i = -xMax + i dx, k in FFT order (src/operators.cu:336-351, src/init.cu:140-162)
    dx = xmax / (n >> 1)
    x = np.empty(n)
    k = np.empty(n)
    dp = np.pi / xmax
    pmax = (np.pi / xmax) * (n >> 1)
    for i in range(n // 2):
        x[i] = -xmax + i * dx
        x[i + n // 2] = i * dx
        k[i] = i * dp
        k[i + n // 2] = -pmax + i * dp
    return x, k, dx


def fft_vectors():
    rng = np.random.default_rng(20241202)
    out = {}
    for n in [8, 12, 16, 30, 32, 64, 97, 128, 192, 256, 384, 512]:
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"in_{n}"] = a
        out[f"fwd_{n}"] = np.fft.fft(a)
        out[f"inv_{n}"] = np.fft.ifft(a) * n   # unnormalised inverse (CUFFT_INVERSE)
    a = rng.standard_normal((8, 16, 32)) + 1j * rng.standard_normal((8, 16, 32))
    out["in_nd"] = a
    out["fwd_nd"] = np.fft.fftn(a)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


def init_psi(shape, grids, seed, phase):
    rng = np.random.default_rng(seed)
    mesh = np.meshgrid(*grids[::-1], indexing="ij")
    r2 = sum(m ** 2 for m in mesh)
    psi = np.exp(-r2 / 2) * (1 + 0.01 * rng.uniform(-1, 1, shape))
    if phase:
        psi = psi * np.exp(1j * rng.uniform(0, 2 * np.pi, shape))
    return psi.astype(np.complex128)


def gpe2d_matlab():
    n, xmax, g, Dt, steps = 64, 8.0, 50.0, 1e-3, 50
    x, k, dx = gpue_grid(n, xmax)
    Y, X = np.meshgrid(x, x, indexing="ij")
    KY, KX = np.meshgrid(k, k, indexing="ij")
    V = 0.5 * (X ** 2 + Y ** 2)
    K = 0.5 * (KX ** 2 + KY ** 2)
    psi0 = init_psi((n, n), [x, x], 7, False)
    psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dx * dx)
    GVr = np.exp(-(V / 1.0) * (Dt / 2))
    GKp = np.exp(-(K / 1.0) * Dt)
    UhbarDt = g / 1.0 * (Dt / 2)
    wfc = psi0.copy()
    hist = {}
    for a in range(1, steps + 1):
        wfc = GVr * np.exp(-np.abs(wfc) ** 2 * UhbarDt) * wfc
        wfc = np.fft.ifft2(GKp * np.fft.fft2(wfc))
        wfc = GVr * np.exp(-np.abs(wfc) ** 2 * UhbarDt) * wfc
        wfc = wfc / (np.sum(np.abs(wfc) ** 2) * dx * dx) ** 0.5
        if a in (1, steps):
            hist[f"psi_{a}"] = wfc.copy()
    np.savez_compressed(os.path.join(HERE, "gpe2d_matlab.npz"), n=n, xmax=xmax, g=g, Dt=Dt, steps=steps,
                        psi0=psi0, **hist)


def gpue_gauge(dim):
    if dim == 2:
        n, xmax, g, omega, steps = 48, 8.0, 60.0, 0.6, 12
    else:
        n, xmax, g, omega, steps = 16, 6.0, 40.0, 0.5, 6
    gdt, dt = 1e-3, 5e-4
    x, k, dx = gpue_grid(n, xmax)
    grids = [x] * dim
    if dim == 2:
        Y, X = np.meshgrid(x, x, indexing="ij")
        KY, KX = np.meshgrid(k, k, indexing="ij")
        Z = KZ = 0.0 * X
    else:
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    oz = 2.0 if dim == 3 else 0.0
    Ax = -Y * omega
    Ay = X * omega
    V = 0.5 * ((X ** 2 + Y ** 2 + (oz * Z) ** 2) + (Ax ** 2 + Ay ** 2))
    K = 0.5 * (KX ** 2 + KY ** 2 + KZ ** 2)
    psi0 = init_psi(X.shape, grids, 11, True)
    psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dx ** dim)
    ax_x, ax_y = dim - 1, dim - 2     # numpy axes of x and y

    def step(w, i, gstate):
        D = gdt if gstate == 0 else dt
        if gstate == 0:
            def op(arg):
                return np.exp(-arg)
        else:
            def op(arg):
                return np.exp(-1j * arg)
        half = lambda w: op(V * D / 2 + g * np.abs(w) ** 2 * (0.5 * D)) * w
        w = half(w)
        w = np.fft.ifftn(op(K * D) * np.fft.fftn(w))
        w = half(w)
        gx = lambda w: np.fft.ifft(op(Ax * KX * D) * np.fft.fft(w, axis=ax_x), axis=ax_x)
        gy = lambda w: np.fft.ifft(op(Ay * KY * D) * np.fft.fft(w, axis=ax_y), axis=ax_y)
        if i % 2 == 1:
            w = gy(gx(w))
        else:
            w = gx(gy(w))
        if gstate == 0:
            w = w / np.sqrt(np.sum(np.abs(w) ** 2) * dx ** dim)
        return w

    w = psi0.copy()
    for i in range(steps):
        w = step(w, i, 0)
    imag = w.copy()
    for i in range(steps):
        w = step(w, i, 1)
    np.savez_compressed(os.path.join(HERE, f"gpue_gauge_{dim}d.npz"), n=n, xmax=xmax, g=g, omega=omega,
                        steps=steps, gdt=gdt, dt=dt, psi0=psi0, psi_imag=imag, psi_real=w)


def gpue_gauge_box():
    """GPUE's step with the per-axis gauge operators on grids whose axes differ
    (the generalisation of src/evolution.cu:163-257, 727-870 to rectangular
    boxes) and on a 1D line, imaginary then real time, in numpy."""
    cases = {
        "line": dict(dims=[128], maxes=[8.0], g=10.0, omega=0.3, gammaY=1.0, oz=0.0, steps=10, seed=21),
        "rect": dict(dims=[64, 32], maxes=[8.0, 5.0], g=40.0, omega=0.45, gammaY=1.2, oz=0.0, steps=10, seed=22),
        "box": dict(dims=[32, 16, 16], maxes=[6.0, 4.0, 5.0], g=30.0, omega=0.4, gammaY=0.9, oz=1.5, steps=6,
                    seed=23),
    }
    gdt, dt = 1e-3, 5e-4
    out = {}
    for name, c in cases.items():
        dim = len(c["dims"])
        ax = [gpue_grid(n, m) for n, m in zip(c["dims"], c["maxes"])]
        while len(ax) < 3:
            ax.append((np.zeros(1), np.zeros(1), 1.0))   # an unused axis: one point at 0
        (x, kx, dx), (y, ky, dy), (z, kz, dz) = ax
        # numpy axes (z, y, x); length-1 axes for the unused ones
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
        om = c["omega"]
        Ax, Ay = -Y * om, X * om
        V = 0.5 * (X ** 2 + (c["gammaY"] * Y) ** 2 + (c["oz"] * Z) ** 2 + Ax ** 2 + Ay ** 2)
        K = 0.5 * (KX ** 2 + KY ** 2 + KZ ** 2)
        dV = dx * (dy if dim > 1 else 1.0) * (dz if dim > 2 else 1.0)
        rng = np.random.default_rng(c["seed"])
        psi0 = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2) * (1 + 0.01 * rng.uniform(-1, 1, X.shape))
        psi0 = psi0 * np.exp(1j * rng.uniform(0, 2 * np.pi, X.shape))
        psi0 = psi0.astype(np.complex128)
        psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dV)
        g = c["g"]

        def step(w, i, gstate):
            D = gdt if gstate == 0 else dt
            op = (lambda a: np.exp(-a)) if gstate == 0 else (lambda a: np.exp(-1j * a))
            half = lambda w: op(V * D / 2 + g * np.abs(w) ** 2 * (0.5 * D)) * w
            w = half(w)
            w = np.fft.ifftn(op(K * D) * np.fft.fftn(w))
            w = half(w)
            gx = lambda w: np.fft.ifft(op(Ax * KX * D) * np.fft.fft(w, axis=2), axis=2)
            gy = lambda w: np.fft.ifft(op(Ay * KY * D) * np.fft.fft(w, axis=1), axis=1)
            if dim == 1:
                w = gx(w)
            elif i % 2 == 1:
                w = gy(gx(w))          # (the z operator is the identity: A_z = 0)
            else:
                w = gx(gy(w))
            if gstate == 0:
                w = w / np.sqrt(np.sum(np.abs(w) ** 2) * dV)
            return w

        w = psi0.copy()
        for i in range(c["steps"]):
            w = step(w, i, 0)
        imag = w.copy()
        for i in range(c["steps"]):
            w = step(w, i, 1)
        shape = tuple(c["dims"][::-1])
        out[f"{name}_dims"] = np.array(c["dims"])
        out[f"{name}_maxes"] = np.array(c["maxes"])
        out[f"{name}_par"] = np.array([g, om, c["gammaY"], c["oz"], c["steps"], gdt, dt])
        out[f"{name}_psi0"] = psi0.reshape(shape)
        out[f"{name}_imag"] = imag.reshape(shape)
        out[f"{name}_real"] = w.reshape(shape)
    np.savez_compressed(os.path.join(HERE, "gpue_gauge_box.npz"), **out)


def vortex_cases():
    """Unit-modulus fields with vortices centred inside chosen plaquettes.

    A plaquette (i, j) spans rows i..i+1 and columns j..j+1; a vortex of charge
    q centred strictly inside it gives that plaquette a phase sum of 2 pi q and
    every other plaquette 0.  Expected markers follow GPUE's rules: value
    |q| for either sign (the sign is lost), in-mask plaquettes only, and a hit
    disables the next in-mask plaquette of the raster stream.
    """
    n, xmax = 64, 8.0
    x, _, dx = gpue_grid(n, xmax)
    cases = {}

    def field(vortices):
        I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        ph = np.zeros((n, n))
        for (ci, cj, q) in vortices:
            ph += q * np.arctan2(I - (ci + 0.5), J - (cj + 0.5))
        return np.exp(1j * ph) * np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / 200.0)

    # 1: isolated vortices of both signs
    v1 = [(20, 20, 1), (20, 40, -1), (40, 30, -1), (45, 45, 1)]
    exp1 = np.zeros((n, n), np.int32)
    for (ci, cj, q) in v1:
        exp1[ci, cj] = abs(q)
    cases["v1"] = (v1, exp1, 100.0)
    # 2: two vortices in raster-adjacent plaquettes: the second is skipped
    v2 = [(30, 30, 1), (30, 31, 1), (30, 40, 1), (30, 42, 1)]
    exp2 = np.zeros((n, n), np.int32)
    exp2[30, 30] = 1
    exp2[30, 40] = 1
    exp2[30, 42] = 1
    cases["v2"] = (v2, exp2, 100.0)
    # 3: a mask excludes the outer vortex: plaquette (i, j) is tested with x[i], x[j]
    v3 = [(32, 32, 1), (5, 5, 1)]
    exp3 = np.zeros((n, n), np.int32)
    exp3[32, 32] = 1
    cases["v3"] = (v3, exp3, 4.0)
    out = {}
    for name, (vs, expm, radius) in cases.items():
        out[f"{name}_psi"] = field(vs)
        out[f"{name}_marker"] = expm
        out[f"{name}_radius"] = radius
    out["x"] = x
    np.savez_compressed(os.path.join(HERE, "vortex_cases.npz"), **out)


if __name__ == "__main__":
    fft_vectors()
    gpe2d_matlab()
    gpue_gauge(2)
    gpue_gauge(3)
    gpue_gauge_box()
    vortex_cases()
    print("golden fixtures written to", HERE)
