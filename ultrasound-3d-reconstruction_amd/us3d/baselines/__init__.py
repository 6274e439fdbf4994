"""Drop-in adapters with the reference's baselines/<method>/render.py signatures."""
