"""Signal plumbing the solver core reports through (status, progress)."""
