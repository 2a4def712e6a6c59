"""The learned sub-grid face closure: a small network that corrects a face's conveyance and crest height from the
face's sub-grid statistics and the water levels either side, trained on coarse-grained fine-grid runs of this same
solver, with the plain table path as its fallback and yardstick."""
