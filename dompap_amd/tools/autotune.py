"""GPU autotuner with the interface of dompap/tools/autotune.py:4-178.

The reference times its neighbour-list methods, scans the skin upward from
``smallest_skin`` in ``step_skin`` increments until a run of ``steps`` steps
gets slower, and times its force methods.  Here every neighbour-list and
force method string runs the same GPU kernels, so the method choice keeps
the simulation's strings; the skin scan is the reference's; and the kernel
shape (lanes per particle, fused or split step) is timed at the chosen skin
and stored in the simulation (``Simulation.set_kernel_tuning``).
"""
from __future__ import annotations

from time import perf_counter


def _time_run(sim, steps):
    tic = perf_counter()
    sim.run(steps)  # returns after the device finished the steps
    return perf_counter() - tic


def autotune(sim, steps=100, test_double_loop=True, smallest_skin=0.1, step_skin=0.1,
             verbose=False, plot=False):
    """Tune skin and kernel shape of ``sim`` (dompap/tools/autotune.py:4-6);
    returns ``sim``."""
    sim_copy = sim.copy()
    sim_copy.step()  # initialise (compile, first list)

    # skin scan (autotune.py:62-90): upward until a run gets slower
    largest_allowed_skin = min(sim.box_vectors) / 2
    skins, times, updates = [], [], []
    skin = smallest_skin
    old_time = float("inf")
    while skin < largest_allowed_skin:
        n0 = sim_copy.number_of_neighbor_list_updates
        sim_copy.set_neighbor_list(skin=skin)
        sim_copy.step()  # the new list and plan, outside the timed run
        t = _time_run(sim_copy, steps)
        skins.append(skin)
        times.append(t)
        updates.append(sim_copy.number_of_neighbor_list_updates - n0)
        skin += step_skin
        if t > old_time:
            break
        old_time = t
    if verbose:
        print("Skin | Time per step (ms) | steps/updates")
        for sk, t, u in zip(skins, times, updates):
            print(f"{sk:4.1f} | {t / steps * 1000:6.4f}             | {steps}/{u} = {steps / max(u, 1):0.1f}")
    best = times.index(min(times))
    skin = skins[best]
    sim.set_neighbor_list(skin=skin)

    # kernel shape at the chosen skin (no counterpart in the reference)
    sim_copy = sim.copy()
    sim_copy.step()
    shapes = [(tpp, 0, fused) for tpp in (1, 2, 4, 8) for fused in (0, 1)]
    shape_times = {}
    for shp in shapes:
        sim_copy.set_kernel_tuning(*shp)
        sim_copy.step()
        shape_times[shp] = _time_run(sim_copy, steps)
        if verbose:
            print(f"threads/particle {shp[0]}, {'fused' if shp[2] else 'split'} step: "
                  f"{shape_times[shp] / steps * 1000:.4f} ms per step")
    best_shape = min(shape_times, key=shape_times.get)
    sim.set_kernel_tuning(*best_shape)
    if verbose:
        print(f"Optimal parameters: skin={skin}, threads per particle={best_shape[0]}, "
              f"{'fused' if best_shape[2] else 'split'} step")
    if plot:
        import matplotlib.pyplot as plt
        plt.plot(skins, times, "o", label="Neighbour list")
        plt.plot(skin, times[best], "ro", label="Fastest time")
        plt.xlabel("Skin")
        plt.ylabel("Time")
        plt.legend()
        plt.show()
    return sim
