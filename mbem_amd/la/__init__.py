"""Linear algebra of the dense solve (mbem's src/mbem/la, dense-path subset)."""
