"""Shared helpers for the path-tracer tests (the reference's tests/pathtrace.rs:49-142)."""
import numpy as np

import oxpt


def psnr(a, b) -> float:
    """LDR PSNR over RGB of two RGBA8 images: 10 log10(255^2 / MSE)."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    mse = float((d * d).mean())
    return float("inf") if mse == 0 else 10.0 * np.log10(255.0 ** 2 / mse)


def mean_rgb(img, x0=0, y0=0, x1=None, y1=None):
    """Mean per channel and standard error of the luminance-ish mean (tests/pathtrace.rs:49-70)."""
    h, w = img.shape[:2]
    x1 = w if x1 is None else x1
    y1 = h if y1 is None else y1
    r = img[y0:y1, x0:x1, :3].astype(np.float64).reshape(-1, 3)
    m = r.mean(axis=0)
    l = r.mean(axis=1)
    se = l.std() / np.sqrt(len(l))
    return m, se


def resolve_params(opts: "oxpt.RenderOptions", samples: int, hdr: bool) -> "oxpt.ResolveParams":
    tm = oxpt.TONE_MAPS[opts.tone_map] if isinstance(opts.tone_map, str) else opts.tone_map
    return oxpt.ResolveParams((1, 3, tm, 1 if hdr else 0), (max(opts.exposure, 0.0), float(samples),
                                                             float(opts.width), 0.0),
                              tuple(c / 255.0 for c in opts.background))


def oracle_frame(scene, opts, env=None, flags=0, threads=0, spp=None):
    """CPU oracle render of the whole frame: (accum, rgba8, hdr)."""
    import oracle

    pk = oxpt.Packed(scene, opts, env)
    osc = oracle.Scene.from_packed(pk)
    n = opts.samples_per_pixel if spp is None else spp
    acc = osc.render(0, n, threads=threads, flags=flags)
    return acc, oracle.resolve(acc, resolve_params(opts, n, False)), oracle.resolve(acc, resolve_params(opts, n, True))


def parity(name, gpu_hdr, gpu_img, cpu_hdr, cpu_img, min_psnr):
    """Statistical parity (tests/pathtrace.rs:88-142): image and 4x4 region means within 4 SE + 1 %, PSNR floor."""
    h, w = gpu_hdr.shape[:2]
    gm, _ = mean_rgb(gpu_hdr)
    cm, cse = mean_rgb(cpu_hdr)
    p = psnr(gpu_img, cpu_img)
    for k in range(3):
        tol = 4.0 * cse + 0.01 * abs(cm[k]) + 1e-4
        assert abs(gm[k] - cm[k]) <= tol, f"{name} ch{k}: {gm[k]} vs {cm[k]} (tol {tol})"
    for ry in range(4):
        for rx in range(4):
            x0, x1, y0, y1 = rx * w // 4, (rx + 1) * w // 4, ry * h // 4, (ry + 1) * h // 4
            a, _ = mean_rgb(gpu_hdr, x0, y0, x1, y1)
            b, se = mean_rgb(cpu_hdr, x0, y0, x1, y1)
            for k in range(3):
                tol = 4.0 * se + 0.01 * abs(b[k]) + 1e-3
                assert abs(a[k] - b[k]) <= tol, f"{name} region ({rx},{ry}) ch{k}: {a[k]} vs {b[k]} (tol {tol})"
    assert p >= min_psnr, f"{name}: PSNR {p:.2f} dB < {min_psnr} dB"
    return p
