"""Progress lines from long tests, written to the terminal past pytest's output capture, so a
run of several minutes reports what it is doing while it runs."""
import concurrent.futures
import time

_config = None


def bind(config):
    global _config
    _config = config


def say(msg: str) -> None:
    pm = _config.pluginmanager if _config is not None else None
    tr = pm.getplugin("terminalreporter") if pm is not None else None
    cap = pm.getplugin("capturemanager") if pm is not None else None
    if tr is None or cap is None:
        print(msg, flush=True)
        return
    # the terminal's descriptor is redirected while a test runs: write with capture suspended
    with cap.global_and_fixture_disabled():
        tr.write_line(msg)
        tr.flush()


def wait(fut, what: str, every: float = 30.0):
    """fut.result(), with a progress line every `every` seconds while it is pending."""
    t0 = time.time()
    while True:
        try:
            return fut.result(timeout=every)
        except concurrent.futures.TimeoutError:
            say(f"[{what}] waiting for the oracle, {time.time() - t0:.0f}s")
