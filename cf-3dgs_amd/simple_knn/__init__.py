"""Drop-in `simple_knn` for MI355X: ``from simple_knn._C import distCUDA2``
(scene/gaussian_model_cf.py:20, called at :194-196)."""
