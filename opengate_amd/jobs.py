"""Split a dose campaign into jobs and merge their outputs (the part of opengate/jobs.py
that the dose path needs: jobs_split assigns primaries to jobs, jobs_merge sums their
image outputs, jobs.py:1083-1111, 3113-3131).

A job is an ordinary Simulation run with `first_history` set: jobs that share one
random_seed and take disjoint, contiguous ranges of history ids draw exactly the random
streams of one long run (history h always uses the Philox counter of h), so merged job
outputs equal the single run's up to the order of FP64 additions. Each image output is
written with its sample-count sidecar (`*_samples.<ext>.json`, dataitems.py:235-245,
1336-1354); merging sums the images and the sample counts, as
ItkImageDataItem.inplace_merge_with does (dataitems.py:1236-1247), and recomputes the
uncertainty images from the merged sums (get_variance_or_uncertainty,
dataitems.py:1956-2005).
"""
from pathlib import Path

import numpy as np

from .actors.dose import insert_suffix_before_extension, read_sample_count, variance_or_uncertainty, write_sample_count
from .image import Image, read_image, write_image

SUMMED = ("edep", "edep_squared", "dose", "dose_squared", "counts")
DERIVED = {"edep_uncertainty": ("edep", "edep_squared"), "dose_uncertainty": ("dose", "dose_squared")}


def split_histories(n_total, n_jobs):
    """[(first_history, n)] of n_jobs contiguous ranges covering n_total histories (sizes
    differ by at most one)."""
    n_total, n_jobs = int(n_total), int(n_jobs)
    if n_jobs < 1 or n_total < 0:
        raise ValueError("n_jobs >= 1 and n_total >= 0 are required")
    base, extra = divmod(n_total, n_jobs)
    out, first = [], 0
    for j in range(n_jobs):
        n = base + (1 if j < extra else 0)
        out.append((first, n))
        first += n
    return out


def merge_dose_outputs(job_dirs, output_filename, output_dir):
    """Merge the DoseActor outputs named output_filename (e.g. "dose.mhd") found in every
    job directory; writes the merged images (and their sidecars) into output_dir and
    returns ({suffix: Image}, number_of_samples)."""
    job_dirs = [Path(d) for d in job_dirs]
    if not job_dirs:
        raise ValueError("no job directories")
    out_dir = Path(output_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    merged, n_samples = {}, None
    for suffix in SUMMED:
        paths = [insert_suffix_before_extension(d / output_filename, suffix) for d in job_dirs]
        present = [p.exists() for p in paths]
        if not any(present):
            continue
        if not all(present):
            missing = [str(p) for p, ok in zip(paths, present) if not ok]
            raise FileNotFoundError(f"output '{suffix}' is missing in some jobs: {missing}")
        total, n = None, 0
        for p in paths:
            img = read_image(p)
            k = read_sample_count(p)
            if k is None:
                raise FileNotFoundError(f"{p}: no sample-count sidecar ({p.name} has no number_of_samples)")
            if total is None:
                total = Image(img.array.astype(np.float64), img.spacing, img.origin, img.direction)
            else:
                if img.array.shape != total.array.shape or not np.allclose(img.spacing, total.spacing) \
                        or not np.allclose(img.origin, total.origin):
                    raise ValueError(f"{p}: image geometry differs between jobs")
                total.array += img.array
            n += k
        if n_samples is None:
            n_samples = n
        elif n != n_samples:
            raise ValueError(f"output '{suffix}': {n} samples, other outputs have {n_samples}")
        merged[suffix] = total
    if not merged:
        raise FileNotFoundError(f"no output of {output_filename} in the job directories")
    for name, (v, v2) in DERIVED.items():
        if v in merged and v2 in merged:
            a = merged[v]
            merged[name] = Image(variance_or_uncertainty(a.array, merged[v2].array, n_samples, "uncertainty"),
                                 a.spacing, a.origin, a.direction)
    for suffix, img in merged.items():
        p = insert_suffix_before_extension(out_dir / output_filename, suffix)
        write_image(img, p)
        write_sample_count(p, n_samples)
    return merged, n_samples
