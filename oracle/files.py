"""Text formats of the variance path, restated from the reference's writers
and readers.

TEST INFRASTRUCTURE ONLY: the checker of libflare_hip's writers and parser.
Nothing in the product imports it.

Writers (the DATE line takes an explicit string so outputs are comparable):
  write_varmap        SparseGP::write_varmap_coefficients (sparse_gp.cpp:
                      1208-1281) with B2::write_to_file (b2.cpp:52-74)
  write_L_inverse     SparseGP::write_L_inverse (sparse_gp.cpp:1283-1339)
  write_sparse        SparseGP::write_sparse_descriptors (sparse_gp.cpp:
                      1341-1396)
C++ stream output in std::scientific with precision 16 is printf "%.16e";
std::fixed with precision 2 (cutoffs) is "%.2f", precision 0 (the kernel
power, NormalizedDotProduct::write_info) is "%.0f".

Readers follow ComputeFlareStdAtom::read_file / read_L_inverse /
read_sparse_descriptors (compute_flare_std_atom.cpp:349-650) with grab
(:652-664): header lines one at a time, value blocks read line by line until
the count is reached.
"""


def _b2_settings(radial, S, N, L, coeff_size, cutoff_fn, cutoffs):
    out = f"{radial}\n{S} {N} {L} {coeff_size}\n{cutoff_fn}\n"
    out += "".join(f"{c:.2f} " for c in cutoffs) + "\n"
    return out


def _hyps(hyps):
    return f"{len(hyps)}\n" + "".join(f"{h:.16e} " for h in hyps) + "\n"


def write_varmap(date, contributor, hyps, S, N, L, radial, cutoff_fn, cutoffs,
                 beta_rows):
    """beta_rows: S rows of F*F values (row-major F x F each)."""
    out = f"DATE: {date} CONTRIBUTOR: {contributor}\n"
    out += _hyps(hyps)
    out += _b2_settings(radial, S, N, L, len(beta_rows[0]), cutoff_fn, cutoffs)
    count = 0
    for row in beta_rows:
        if count != 0:
            out += "\n"
            count = 0
        for v in row:
            out += ("  " if v > 0 else " ") + f"{v:.16e}"
            count += 1
            if count == 5:
                count = 0
                out += "\n"
    return out


def write_L_inverse(date, contributor, power, hyps, S, N, L, radial, cutoff_fn,
                    cutoffs, Linv):
    """Linv: M x M (nested lists or array); the lower triangle is written."""
    M = len(Linv)
    out = f"L_inverse_block file DATE: {date} CONTRIBUTOR: {contributor}\n"
    out += f"{power:.0f}\n"
    out += _hyps(hyps)
    out += _b2_settings(radial, S, N, L, 1, cutoff_fn, cutoffs)
    out += f"{M}\n"
    count = 1
    for j in range(M):
        for k in range(j + 1):
            out += f"{Linv[j][k]:.16e} "
            if count % 5 == 0:
                out += "\n"
            count += 1
    if count % 5 != 0:
        out += "\n"
    return out


def write_sparse(date, contributor, n_by_type, rows, norms):
    """rows: raw sparse descriptors (M x F, species-grouped), norms: M."""
    M = sum(n_by_type)
    out = f"sparse_descriptors file DATE: {date} CONTRIBUTOR: {contributor}\n"
    out += "1\n"
    out += f"0 {M} {len(n_by_type)}\n"
    count = 1
    r = 0
    for c in n_by_type:
        out += f"{c}\n"
        for _ in range(c):
            for x in rows[r]:
                v = 0.0 if norms[r] < 1e-8 else x / norms[r]
                out += f"{v:.16e} "
                if count % 5 == 0:
                    out += "\n"
                count += 1
            r += 1
        if count % 5 != 1:
            out += "\n"
    return out


class _Lines:
    def __init__(self, text):
        self.lines = text.split("\n")
        if self.lines and self.lines[-1] == "":
            self.lines.pop()
        self.i = 0

    def next(self):
        if self.i >= len(self.lines):
            raise ValueError("file ends early")
        line = self.lines[self.i]
        self.i += 1
        return line

    def grab(self, n):
        vals = []
        while len(vals) < n:
            vals.extend(float(t) for t in self.next().split())
        if len(vals) != n:
            raise ValueError("more values on a line than expected")
        return vals


def read_varmap(text):
    """compute_flare_std_atom.cpp:349-451"""
    f = _Lines(text)
    f.next()
    n_hyps = int(f.next().split()[0])
    hyps = [float(t) for t in f.next().split()[:4]]
    radial = f.next().split()[0]
    S, N, L, beta_size = (int(t) for t in f.next().split()[:4])
    cutoff_fn = f.next().split()[0]
    cutoffs = f.grab(S * S)
    nr = S * N
    F = nr * (nr + 1) // 2 * (L + 1)
    if F * F != beta_size:
        raise ValueError("Beta size doesn't match the number of descriptors.")
    beta = f.grab(S * beta_size)
    return dict(n_hyps=n_hyps, hyps=hyps, radial=radial, S=S, N=N, L=L, F=F,
                cutoff_fn=cutoff_fn, cutoffs=cutoffs,
                beta=[beta[s * beta_size:(s + 1) * beta_size] for s in range(S)])


def read_L_inverse(text):
    """compute_flare_std_atom.cpp:453-563"""
    f = _Lines(text)
    f.next()
    power = int(f.next().split()[0])
    n_hyps = int(f.next().split()[0])
    hyps = [float(t) for t in f.next().split()[:4]]
    radial = f.next().split()[0]
    S, N, L, n_kernels = (int(t) for t in f.next().split()[:4])
    cutoff_fn = f.next().split()[0]
    cutoffs = f.grab(S * S)
    M = int(f.next().split()[0])
    vals = f.grab(M * (M + 1) // 2)
    Linv = [[0.0] * M for _ in range(M)]
    c = 0
    for j in range(M):
        for k in range(j + 1):
            Linv[j][k] = vals[c]
            c += 1
    return dict(power=power, n_hyps=n_hyps, hyps=hyps, radial=radial, S=S, N=N,
                L=L, n_kernels=n_kernels, cutoff_fn=cutoff_fn, cutoffs=cutoffs,
                M=M, Linv=Linv)


def read_sparse(text, n_kernels, n_clusters, F):
    """compute_flare_std_atom.cpp:565-650"""
    f = _Lines(text)
    f.next()
    if int(f.next().split()[0]) != n_kernels:
        raise ValueError("n_kernals in sparse_descriptors and L_inv not match")
    _, n_clst, n_types = (int(t) for t in f.next().split()[:3])
    if n_clst != n_clusters:
        raise ValueError("n_clusters in sparse_descriptors and L_inv not match")
    n_by_type, rows = [], []
    for _ in range(n_types):
        # a species with no clusters after a part-filled line leaves a blank
        # line in the writer's output, which the reference reads as the next
        # count (sscanf fails, the count is indeterminate); skip it instead
        line = f.next()
        while not line.strip():
            line = f.next()
        c = int(line.split()[0])
        n_by_type.append(c)
        vals = f.grab(c * F)
        rows.extend(vals[j * F:(j + 1) * F] for j in range(c))
    return dict(n_types=n_types, n_by_type=n_by_type, rows=rows)
