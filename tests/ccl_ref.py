"""CPU reference for particle labelling: scipy.ndimage.label (6-neighbour
structure) on the box, then labels that touch across a periodic face are
merged with a small union-find.  Returns component sizes, largest first."""
import numpy as np
from scipy import ndimage


def component_sizes(mask, periodic=True):
    lab, n = ndimage.label(mask, structure=ndimage.generate_binary_structure(3, 1))
    parent = np.arange(n + 1)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    if periodic:
        for ax in range(3):
            first = np.take(lab, 0, axis=ax)
            last = np.take(lab, lab.shape[ax] - 1, axis=ax)
            both = (first > 0) & (last > 0)
            for a, b in zip(first[both], last[both]):
                ra, rb = find(a), find(b)
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
    roots = np.array([find(i) for i in range(n + 1)])
    counts = np.bincount(roots[lab.ravel()], minlength=n + 1)
    counts[0] = 0
    sizes = counts[counts > 0]
    return np.sort(sizes)[::-1]


def mean_radius(sizes):
    return float(np.mean(np.cbrt(3.0 * sizes / (4.0 * np.pi)))) if len(sizes) else 0.0
