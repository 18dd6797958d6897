"""Checkpoints of an engine's dynamic state: per replica the Gear
derivatives y0..y5, the generator state, the time, the QA contact set with
its sync budgets and the LJ attractive wall pairs, saved to one .npz (no
pickled objects).

The reference keeps exactly these in its store as saved variables
(xmd/src/freeform/main.cpp:143-253,655-672; store.h:147-184) and rebuilds
its neighbour lists after loading them (nl.invalid starts true,
main.cpp:374); restoring through xmd_set_state does the same, so a run
restarted from a checkpoint continues as the saved run would after a
forced list rebuild at that step.
"""
import numpy as np

FORMAT = 1


def save(engine, path):
    """Write every replica's state of `engine` (Engine or oracle driver)."""
    arrs = {"format": np.array([FORMAT]), "n_replicas": np.array([engine.nrep])}
    for r in range(engine.nrep):
        y, _, rng, t = engine.get_state(r)
        cont, sync = engine.get_qa_state(r)
        arrs[f"y{r}"] = y
        arrs[f"rng{r}"] = np.array([rng], dtype=np.uint64)
        arrs[f"t{r}"] = np.array([t])
        arrs[f"qa{r}"] = cont
        arrs[f"sync{r}"] = sync
        if len(engine.system.lj_walls):
            st, jo, rt = engine.get_wall_state(r)
            arrs[f"wall_status{r}"], arrs[f"wall_joint{r}"], arrs[f"wall_ref_t{r}"] = st, jo, rt
    np.savez(path, **arrs)


def load(engine, path):
    """Restore a checkpoint written by save() into an engine built from the
    same parameters and topology."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["format"][0]) != FORMAT:
            raise ValueError(f"unknown checkpoint format {int(z['format'][0])}")
        if int(z["n_replicas"][0]) != engine.nrep:
            raise ValueError("checkpoint replica count differs from the engine's")
        for r in range(engine.nrep):
            engine.set_state(r, z[f"y{r}"], int(z[f"rng{r}"][0]), float(z[f"t{r}"][0]))
            engine.set_qa_state(r, z[f"qa{r}"], z[f"sync{r}"])
            if f"wall_status{r}" in z:
                engine.set_wall_state(r, z[f"wall_status{r}"], z[f"wall_joint{r}"], z[f"wall_ref_t{r}"])
