"""Training-side NEP featurisation of site clouds (05/NEP.py, 04/NEP.py call surface).

NEP(pos) takes pair site clouds pos (Np, 2*npts, 3); get_g_fast() returns the
(Np, D) float32 descriptors [g_rad | g_ang] with D = (nrad+1) + (nang+1)*lmax,
each cloud about its own centroid, radial terms weighted per site by the type
weights of set_pts_types (05/NEP.py:152-163: types+1, square-rooted when the
largest exceeds 4).  The work runs in the HIP cloud-descriptor kernel
(nepx_cloud_descriptor) in chunks; there is no host path.

Differences from the reference, by design: get_g() (the reference's FP64 host
loop, 05/NEP.py:222-283, which ignores the type weights) also runs on the GPU
in FP32 without weights; hypers with no kernel instance raise NotImplementedError.
"""
import numpy as np
import torch

from . import ops


def calculate_cutoff(pos):
    """Largest site distance from its cloud centroid over all clouds, rounded up to 0.01
    (05/NEP.py:128-134)."""
    pos = np.asarray(pos)
    c = pos.mean(axis=1)
    r = np.linalg.norm(pos - c[:, None, :], axis=-1)
    return np.ceil(r.max() * 100) / 100


class NEP:
    def __init__(self, pos, device_id=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float32)
        self.npts = self.pos.shape[1] // 2
        self.types = None
        self.device = torch.device("cuda", 0 if device_id is None else int(device_id))
        self.chunk = 1 << 20

    def set_npts(self, n):
        self.npts = int(n)

    def set_device(self, device_id):
        self.device = torch.device("cuda", int(device_id))

    def set_pts_types(self, hparams):
        if "pts_types" in hparams:
            t = np.asarray(hparams["pts_types"], dtype=np.float64) + 1.0
        else:
            t = np.ones(self.npts * 2)
        if t.max() > 4:
            t = np.sqrt(t)
        self.types = t.astype(np.float32)

    def set_hypers(self, hypers):
        self.nrad, self.nang, self.lmax = int(hypers[0]), int(hypers[1]), int(hypers[2])
        self.cutoff = float(hypers[3])

    def _run(self, weights):
        if self.pos.shape[1] != 2 * self.npts:
            raise ValueError(f"clouds have {self.pos.shape[1]} sites, expected 2*npts = {2 * self.npts}")
        out = []
        for a in range(0, len(self.pos), self.chunk):
            x = torch.from_numpy(self.pos[a:a + self.chunk]).to(self.device)
            try:
                G = ops.cloud_descriptor(x, self.nrad, self.nang, self.lmax, self.cutoff, weights)
            except Exception as e:
                if "no cloud descriptor kernel" in str(e):
                    raise NotImplementedError(str(e)) from e
                raise
            out.append(G.cpu().numpy())
        return np.concatenate(out) if out else np.zeros((0, (self.nrad + 1) + (self.nang + 1) * self.lmax),
                                                        np.float32)

    def get_g_fast(self):
        return self._run(self.types)

    def get_g(self):
        return self._run(None)


def pts_to_nep(pts_all, hparams, device_id=0, nep_cutoff=None):
    """Featurise all clouds (05/Processors.py:46-92): returns (G float32, nep_cutoff)."""
    hyp = hparams["nep_hyper"] if isinstance(hparams, dict) else hparams
    if nep_cutoff is None:
        nep_cutoff = calculate_cutoff(pts_all)
    nep = NEP(pts_all, device_id)
    nep.set_hypers([int(hyp[0]), int(hyp[1]), int(hyp[2]), nep_cutoff])
    nep.set_npts(np.asarray(pts_all).shape[1] // 2)
    nep.set_pts_types(hparams if isinstance(hparams, dict) else {})
    return nep.get_g_fast(), nep_cutoff
