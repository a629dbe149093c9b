"""Per-call time limit for the first execution of an implementation.

Same mechanism as npbench/infrastructure/timeout_decorator.py:46-72: a
``threading.Timer`` calls ``_thread.interrupt_main()``, which raises
KeyboardInterrupt in the main thread when it next runs Python bytecode.  A
blocking native call is not interrupted, so the MI355X adapter waits for the
GPU in short polls (npbench_amd.device.sync) to give the interrupt a chance.
"""

import _thread
import sys
import threading


def _expire(name):
    print(f"{name} took too long", file=sys.stderr, flush=True)
    _thread.interrupt_main()


def exit_after(seconds):
    """Decorator: raise KeyboardInterrupt in the main thread after ``seconds``."""

    def wrap(fn):
        def inner(*args, **kwargs):
            timer = threading.Timer(seconds, _expire, args=[fn.__name__])
            timer.daemon = True
            timer.start()
            try:
                return fn(*args, **kwargs)
            finally:
                timer.cancel()

        inner.__name__ = fn.__name__
        return inner

    return wrap
