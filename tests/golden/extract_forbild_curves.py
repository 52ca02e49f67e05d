"""Write forbild_cgls_curves.json: the PSNR / SSIM curves the reference's
notebook printed for its CGLS reconstructions of the FORBILD phantom
(python/main copy 2.ipynb, cell 27: B0, cell 28: B1; pylops
regularized_inversion with niter = 2, 4, ..., 64 at 512 x 512 pixels, 900
views, 855 bins).  Reads the notebook's saved outputs as data; runs nothing
from the reference.  Run in a checkout that has the reference at
/root/reference."""
import json
import os
import re
import sys

NB = "/root/reference/python/main copy 2.ipynb"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "forbild_cgls_curves.json")


def main():
    cells = json.load(open(NB))["cells"]
    out = {"source": "python/main copy 2.ipynb cells 27 (B0) and 28 (B1), printed outputs"}
    for basis, idx in (("B0", 27), ("B1", 28)):
        text = "".join("".join(o.get("text", "")) for o in cells[idx].get("outputs", []))
        rows = re.findall(r"^(\d+) \tPSNR: ([\d.]+) \tRMSE: [\d.e-]+ \tFWHM: \S+ \tSSIM: ([\d.]+)", text, re.M)
        out[basis] = [dict(iters=int(k), psnr=float(p), ssim=float(s)) for k, p, s in rows]
    json.dump(out, open(OUT, "w"), indent=1)
    print("wrote", OUT, {b: len(out[b]) for b in ("B0", "B1")})


if __name__ == "__main__":
    sys.exit(main())
