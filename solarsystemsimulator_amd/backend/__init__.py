"""Drop-in modules with the reference's names (nbody_kernel, selfgrav_kernel,
pm_kernel).  Put this directory on sys.path ahead of the reference's
``backend/`` so ``import selfgrav_kernel`` in film_producer.py / server.py
resolves here."""
