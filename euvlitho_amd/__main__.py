"""Command-line drivers that replace the reference programs and scripts,
reading and writing the same files in the working directory:

  python -m euvlitho_amd m3d     mask.bin -> inputlm.csv, inputxx.csv   (m3d.cpp main)
  python -m euvlitho_amd abbe    mask.bin -> emint.csv | emintY.csv     (intensity.cpp main)
  python -m euvlitho_amd linear  mask.bin + inputxx.csv -> intlinear.csv, intft.csv  (linear.py __main__)
  python -m euvlitho_amd socs    mask.bin + inputxx.csv -> intsocs.csv  (socs.py __main__)
  python -m euvlitho_amd mask    -> mask.bin, maskimage.csv             (mask.cpp main)
  python -m euvlitho_amd predict mask.bin + model/*/model.ckpt -> inputxx.csv  (predict.py __main__)
  python -m euvlitho_amd compress mask.bin + inputxx.csv + inputlm.csv -> maskdata/ ampdata/ inputlm/ (compress.py)
  python -m euvlitho_amd dataset  -> the same npy files, generated on the GPU (mask.cpp -> m3d.cpp -> compress.py)

--preset picks the problem (reference: the drivers' own optics; c1, c2: the
high-NA clips); --format g writes the reference C++ 6-digit numbers.
"""
import argparse
import os
import sys
import time

from . import csvio
from .api import Session
from .config import preset


def write_masks(out, ndiv, ndata, seed, cd, gap):
    """mask.cpp:28-62: ndata line/space clips (drawn on the host, rasterised on
    the GPU) -> mask.bin; the first clip as 0/1 text -> maskimage.csv (row j
    holds mask2d[NDIVX*j + i] for every i, each followed by a comma)."""
    import numpy as np
    from .masks import line_space_bits, unpack_bits
    bits = line_space_bits(ndiv, ndiv, ndata, seed, cd, gap).cpu().numpy()
    with open(out("mask.bin"), "wb") as f:
        f.write(bits.tobytes())
    m = unpack_bits(bits.tobytes(), ndiv, ndiv, 0).reshape(ndiv, ndiv)
    lines = np.where(m == 1, "1,", "0,")
    with open(out("maskimage.csv"), "w") as f:
        f.write("\n".join("".join(row) for row in lines) + "\n")


def predict(out, mask, model_dir, ndiv, nout):
    """predict.py:122-151: the six networks from model_dir/<name>/model.ckpt,
    the factors from re0/factor0.csv, rex/factorx.csv, rey/factory.csv, the
    clips of mask.bin through compmask -> inputxx.csv (np.savetxt blocks)."""
    import numpy as np
    from .cnn import CNNmodel, CNNmodels, compmask, predict_m3d, write_inputxx
    d = lambda *p: os.path.join(model_dir, *p)
    models = [CNNmodel.load_from_checkpoint(d("re0", "model.ckpt")),
              CNNmodel.load_from_checkpoint(d("im0", "model.ckpt"))]
    models += [CNNmodels.load_from_checkpoint(d(n, "model.ckpt")) for n in ("rex", "imx", "rey", "imy")]
    factor0 = np.loadtxt(d("re0", "factor0.csv"), delimiter=",")
    factorx = np.loadtxt(d("rex", "factorx.csv"), delimiter=",")
    factory = np.loadtxt(d("rey", "factory.csv"), delimiter=",")
    x = compmask(mask, ndiv, nout)
    write_inputxx(out("inputxx.csv"), predict_m3d(x, *models, factor0, factorx, factory))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m euvlitho_amd", description=__doc__.split("\n\n")[0])
    ap.add_argument("command", choices=["m3d", "abbe", "linear", "socs", "mask", "predict", "compress", "dataset"])
    ap.add_argument("--preset", default="reference")
    ap.add_argument("--mask", default="mask.bin")
    ap.add_argument("--inputxx", default="inputxx.csv")
    ap.add_argument("--inputlm", default="inputlm.csv", help="compress: the pupil point list")
    ap.add_argument("--ndata", type=int, default=1, help="m3d, mask: clips in mask.bin")
    ap.add_argument("--ndiv", type=int, default=2048, help="mask: clip size in pixels (NDIVX = NDIVY)")
    ap.add_argument("--seed", type=int, default=1, help="mask: srand seed (the reference program's: 1)")
    ap.add_argument("--cd", type=int, default=56, help="mask: column band width")
    ap.add_argument("--gap", type=int, default=80, help="mask: line/space segment unit")
    ap.add_argument("--model-dir", default="../model",
                    help="predict: directory of re0/ im0/ rex/ imx/ rey/ imy/ with model.ckpt and the factor csv files")
    ap.add_argument("--nout", type=int, default=512, help="predict: network input size (compmask output)")
    ap.add_argument("--polar", default="X", choices=["X", "Y"], help="abbe: polarisation")
    ap.add_argument("--z", type=float, default=0.0, help="defocus (nm)")
    ap.add_argument("--format", default="repr", choices=list(csvio.FORMATS))
    ap.add_argument("--out", default=".", help="output directory")
    a = ap.parse_args(argv)

    out = lambda name: os.path.join(a.out, name)
    t0 = time.perf_counter()
    if a.command == "mask":
        write_masks(out, a.ndiv, a.ndata, a.seed, a.cd, a.gap)
        print(f"elapsed time (s) {time.perf_counter() - t0:.3f}", file=sys.stderr)
        return 0
    if a.command == "predict":
        predict(out, a.mask, a.model_dir, a.ndiv, a.nout)
        print(f"elapsed time (s) {time.perf_counter() - t0:.3f}", file=sys.stderr)
        return 0
    if a.command == "compress":
        from .dataset import compress
        compress(a.mask, a.inputxx, a.inputlm, a.ndata, a.out, a.ndiv, a.nout)
        print(f"elapsed time (s) {time.perf_counter() - t0:.3f}", file=sys.stderr)
        return 0
    if a.command == "dataset":
        from .dataset import generate
        generate(a.ndata, a.out, preset(a.preset), a.seed, a.cd, a.gap, a.nout, write_intermediates=True,
                 progress=lambda i, n: print(f"clip {i}/{n}", file=sys.stderr, flush=True))
        print(f"elapsed time (s) {time.perf_counter() - t0:.3f}", file=sys.stderr)
        return 0
    ses = Session(preset(a.preset))
    if a.command == "m3d":
        lin, min_, _ = ses.cfg.pupil
        csvio.write_inputlm(out("inputlm.csv"), lin, min_, ses.xinput)
        for s, (a0, ax, ay) in enumerate(ses.m3d(a.mask, a.ndata)):
            csvio.write_inputxx(out("inputxx.csv"), a0, ax, ay, ses.xinput, a.format, append=s > 0)
    elif a.command == "abbe":
        img = ses.abbe(a.mask, a.polar, a.z)
        csvio.write_image(out("emint.csv" if a.polar == "X" else "emintY.csv"), img, a.format)
    else:
        re0, im0, rex, imx, rey, imy = csvio.read_inputxx(a.inputxx)[0]
        if a.command == "linear":
            il, it = ses.linear(a.z, a.mask, re0, im0, rex, imx, rey, imy)
            csvio.write_image(out("intlinear.csv"), il, a.format)
            csvio.write_image(out("intft.csv"), it, a.format)
        else:
            img = ses.socs(a.z, a.mask, re0 + 1j * im0, rex + 1j * imx, rey + 1j * imy)
            csvio.write_image(out("intsocs.csv"), img, a.format)
    print(f"elapsed time (s) {time.perf_counter() - t0:.3f}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
