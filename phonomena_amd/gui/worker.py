"""``WorkerSignals`` without a hard Qt dependency.

The reference's solver core imports ``gui.worker.WorkerSignals``
(phonomena/gui/worker.py:6-35), a PyQt5 ``QObject`` whose ``status`` and
``progress`` signals the step loop emits into (base_solver.py:235-280).
When PyQt5 is importable this module re-exports that class shape as real Qt
signals; otherwise it provides the same attributes as plain callbacks, so
the solver never needs Qt.
"""
try:  # pragma: no cover - PyQt5 is not installed in the build image
    from PyQt5 import QtCore
    from PyQt5.QtCore import pyqtSignal

    class WorkerSignals(QtCore.QObject):
        finished = pyqtSignal()
        success = pyqtSignal()
        error = pyqtSignal(tuple)
        warning = pyqtSignal(tuple)
        result = pyqtSignal(object)
        progress = pyqtSignal(int)
        status = pyqtSignal(str)
        quit = pyqtSignal()

    HAVE_QT = True
except ImportError:
    HAVE_QT = False

    class Signal:
        """Minimal stand-in for a Qt signal: connect() callbacks, emit() calls them."""

        def __init__(self):
            self._slots = []

        def connect(self, fn):
            self._slots.append(fn)

        def disconnect(self, fn=None):
            if fn is None:
                self._slots.clear()
            else:
                self._slots.remove(fn)

        def emit(self, *args):
            for fn in list(self._slots):
                fn(*args)

    class WorkerSignals:
        """Same signal names as the reference's WorkerSignals."""

        def __init__(self):
            self.finished = Signal()
            self.success = Signal()
            self.error = Signal()
            self.warning = Signal()
            self.result = Signal()
            self.progress = Signal()
            self.status = Signal()
            self.quit = Signal()
