"""integrate() driver, mirroring src/lib.rs:132-187."""
MAX_TIMESTEP = 10_000_000


def integrate(pde, max_time, save_intervall=None):
    """Step `pde` (update/get_time/get_dt/callback/exit) until max_time (src/lib.rs:155-187)."""
    timestep = 0
    eps_dt = pde.get_dt() * 1e-4
    while True:
        pde.update()
        timestep += 1
        if save_intervall is not None:
            t, dt = pde.get_time(), pde.get_dt()
            r = t % save_intervall
            if r < dt / 2.0 or r > save_intervall - dt / 2.0:
                pde.callback()
        if pde.get_time() + eps_dt >= max_time:
            print("time limit reached: %r" % pde.get_time())
            break
        if timestep >= MAX_TIMESTEP:
            print("timestep limit reached: %r" % timestep)
            break
        if pde.exit():
            print("break criteria triggered")
            break
